// One-shot LDS box kernel for star stencils (gfx950).
//
// The plane-sweeping tile kernel (stencil_tile.h) reads every value once per
// sweep, but a block's planes are a chain of dependent steps: one memory
// round trip and one barrier per plane.  On grids that sit in the caches
// (the multigrid levels of a Poisson solve: 95^3 is 7 MB per array) that
// chain, not bandwidth, sets the time.  Here a block of TZ x NTY threads owns
// a TZ x NTY tile of the (i1, i2) plane and Q consecutive i0 planes, and
// issues every load it needs at once:
//   * its own column over the Q planes and R planes on either side (the
//     axis-0 neighbours, kept in registers);
//   * the in-plane halo (R-wide rows and columns around the tile) of the Q
//     planes, precomputed in-plane offsets as in the tile kernel;
//   * the aux values (rho of a Jacobi sweep or a residual) of the Q planes.
// The Q centre planes and their halos go to LDS, one barrier, then each
// thread evaluates its Q points and stores them.  The column halo costs
// 2R/Q extra reads per point, served by the L2 when the neighbouring chunks
// run on the same XCD (tiles are ordered per XCD as in the tile kernel).
//
// Terms are summed in the tile kernel's order (centre, +-k along axis 2,
// axis 1, axis 0: c/bmgs/fd.c:41-44 with the NewGUCLaplace order) with the
// same epilogue, so results are bit-identical to the tile kernel and to the
// reference C.  Periodic axes wrap, zero-
// boundary axes read zeros, complex data get the Bloch phase of every side
// they wrap across (ghost_fix, the tile kernel's corrections).
#pragma once
#include "stencil_tile.h"

namespace gpaw {

constexpr int BOX_TZ = 32;     // columns (lanes along i2)
constexpr int BOX_NTY = 8;     // rows (one i1 row per thread row)
constexpr int BOX_NT = BOX_TZ * BOX_NTY;

template <class V, int R, int Q>
struct BoxGeom {
    static constexpr int TZ = BOX_TZ, TY = BOX_NTY;
    static constexpr int LW = TZ + 2 * R;                 // LDS row length
    static constexpr int LH = TY + 2 * R;                 // LDS rows
    static constexpr int NHALO = 2 * R * (TZ + TY);       // halo cells per plane
    static constexpr int NH = (NHALO + BOX_NT - 1) / BOX_NT;
    static constexpr int LDS = Q * LH * LW;               // Q plane buffers
};

// FZ (ST_JACOBI, real data): `in` is rho and every loaded value v first
// becomes the sweep from phi = 0 with weight w0, the element-wise
// st_epilogue(0, 0, v) of k_jacobi_from_zero; the kernel then runs the
// second sweep on those values, so two sweeps from zero take one launch and
// give the bits of the two launches.
template <class V, int R, int Q, int MODE, bool NORM, bool WRITE, bool FZ = false>
__global__ __launch_bounds__(BOX_NT) void k_star_box(const V* __restrict__ in, V* __restrict__ out,
                                                     const V* __restrict__ aux, Geom g, Phases ph,
                                                     TileCoefs sc, double w, TileShape sh,
                                                     i64 ng, double* __restrict__ partial,
                                                     double w0)
{
    typedef typename VT<V>::R RT;
    typedef BoxGeom<V, R, Q> BG;
    constexpr int TZ = BG::TZ, TY = BG::TY, LW = BG::LW, LH = BG::LH, NHALO = BG::NHALO,
                  NH = BG::NH;
    __shared__ V lds[BG::LDS];

    const int b = blockIdx.x;
    const int t = xcd_tile(b, sh.nblocks);
    const int tzi = t % sh.ntz;
    int rest = t / sh.ntz;
    const int tyb = rest % sh.nty;
    rest /= sh.nty;
    const int ch = rest % sh.nchunk;
    const int band = rest / sh.nchunk;
    const i64 boff = (i64)band * ng;
    const V* __restrict__ a = in + boff;
    V* __restrict__ o = out + boff;
    // ST_HPSI: one real potential shared by all bands
    const V* __restrict__ sa = (aux && MODE != ST_HPSI) ? aux + boff : aux;

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

    // every load first: the column over planes i0s-R .. i0s+Q-1+R, the halo
    // cells and aux values of planes i0s .. i0s+Q-1.  Planes past the grid
    // (a partial last chunk) load a valid address and are never stored.
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
        if (MODE != ST_APPLY) {
            const bool ok = inr && i0s + p < n0;
            sn[p] = aux_value<MODE>(sa, 0, ok ? (i64)(i0s + p) * plane + pref_off(qr) : 0);
        }
    }
    // wait for every load here: a load consumed only after the barrier could
    // be scheduled behind the stores and cost a second round trip
#pragma unroll
    for (int p = 0; p < Q; p++) {
#pragma unroll
        for (int k = 0; k < NH; k++) vtouch(hq[p][k]);
        if (MODE != ST_APPLY) vtouch(sn[p]);
    }
    auto fz = [&](V v) -> V {
        return FZ ? st_epilogue<ST_JACOBI>(vzero<V>(), vzero<V>(), v, (RT)w0, (RT)sc.c[0]) : v;
    };
#pragma unroll
    for (int p = 0; p < Q + 2 * R; p++) q[p] = fz(ghost_fix(q[p], qf[p], qr, ph));
#pragma unroll
    for (int p = 0; p < Q; p++) {
        V* B = lds + p * (LH * LW);
        B[(R + tyi) * LW + R + tx] = q[R + p];
#pragma unroll
        for (int k = 0; k < NH; k++)
            if (hoff[k] >= 0) B[hoff[k]] = fz(ghost_fix(hq[p][k], qf[R + p], hr[k], ph));
    }
    __syncthreads();

    double acc = 0.0;
#pragma unroll
    for (int p = 0; p < Q; p++) {
        const V* row = lds + p * (LH * LW) + (R + tyi) * LW + R + tx;
        constexpr bool RIN = NORM && MODE == ST_JACOBI;
        const V f = q[R + p];
        V x = vzero<V>(), xr = vzero<V>();
        if (MODE != ST_JACOBI) x = vmac(x, f, (RT)sc.c[0]);
        if (RIN) xr = vmac(xr, f, (RT)sc.c[0]);
#pragma unroll
        for (int k = 1; k <= R; k++) {
            x = vmac(x, row[k], (RT)sc.c[k]);
            if (RIN) xr = vmac(xr, row[k], (RT)sc.c[k]);
        }
#pragma unroll
        for (int k = 1; k <= R; k++) {
            x = vmac(x, row[-k], (RT)sc.c[R + k]);
            if (RIN) xr = vmac(xr, row[-k], (RT)sc.c[R + k]);
        }
#pragma unroll
        for (int k = 1; k <= R; k++) {
            x = vmac(x, row[k * LW], (RT)sc.c[2 * R + k]);
            if (RIN) xr = vmac(xr, row[k * LW], (RT)sc.c[2 * R + k]);
        }
#pragma unroll
        for (int k = 1; k <= R; k++) {
            x = vmac(x, row[-k * LW], (RT)sc.c[3 * R + k]);
            if (RIN) xr = vmac(xr, row[-k * LW], (RT)sc.c[3 * R + k]);
        }
#pragma unroll
        for (int k = 1; k <= R; k++) {
            x = vmac(x, q[R + p + k], (RT)sc.c[4 * R + k]);
            if (RIN) xr = vmac(xr, q[R + p + k], (RT)sc.c[4 * R + k]);
        }
#pragma unroll
        for (int k = 1; k <= R; k++) {
            x = vmac(x, q[R + p - k], (RT)sc.c[5 * R + k]);
            if (RIN) xr = vmac(xr, q[R + p - k], (RT)sc.c[5 * R + k]);
        }
        const V s = MODE == ST_APPLY ? vzero<V>() : sn[p];
        const V r = st_epilogue<MODE>(x, f, s, (RT)w, (RT)sc.c[0]);
        const bool own = inr && i0s + p < n0;
        if (WRITE && own) o[(i64)(i0s + p) * plane + pref_off(qr)] = r;
        if (NORM && own) {
            if (MODE == ST_APPLY)
                acc += tdot(f, r);
            else if (RIN)
                acc += tnorm2(st_epilogue<ST_RESID>(xr, f, s, (RT)w, (RT)sc.c[0]));
            else
                acc += tnorm2(r);
        }
    }
    if (NORM) {
        __shared__ double red[BOX_NT];
        const double sum = block_sum<BOX_NT>(acc, red);
        if (threadIdx.x == 0 && threadIdx.y == 0) partial[b] = sum;
    }
}

// Box shapes: Q planes per block (2, 4 or 8); false if the grid is too small
// for one wrap of the halo or the mode has no box kernel.
bool box_geometry(const Geom& g, int R, int nin, int Q, TileShape& sh);
bool box_mode_ok(int dtype, int mode);
int box_default_q(const Geom& g, int R, int nin);
hipError_t box_launch(const Stencil& s, int dtype, int mode, const void* in, void* out,
                      const void* aux, const Geom& g, const Phases& ph, double w,
                      const TileShape& sh, double* partial, bool write, hipStream_t st);
hipError_t box_fz_f64(const Stencil& s, const void* rho, void* out, const Geom& g, double w0,
                      double w1, const TileShape& sh, hipStream_t st);
hipError_t box_fz_f32(const Stencil& s, const void* rho, void* out, const Geom& g, double w0,
                      double w1, const TileShape& sh, hipStream_t st);

}  // namespace gpaw
