// Finite-difference stencil kernels for gfx950: apply, residual and
// weighted-Jacobi smoothing.
//
// Arithmetic follows c/bmgs/fd.c:16-51 (x = 0; x += a[off_c] * coef_c in
// coefficient order) and c/bmgs/relax.c:47-71 (x over c >= 1, then
// (1-w)*f + w*(s - x)/coef_0), evaluated without FMA contraction.
//
// Two kernels:
//  * k_star: the hot path.  For the canonical orthogonal Laplacian order
//    (centre, +-k along axis 2, +-k along axis 1, +-k along axis 0) each
//    thread owns one (i1, i2) column and sweeps a run of i0 planes, keeping
//    the 2R+1 values along i0 in a register queue, so axis-0 neighbours are
//    read from HBM once.  In-plane neighbours come from the same plane rows
//    that the neighbouring lanes / waves of the block load (L1/L2 hits).
//    Waves are 64 lanes along i2, so every HBM access is a coalesced row.
//  * k_generic: any stencil (non-orthogonal cells, Mehrstellen, gradients):
//    one thread per point, coefficients read through the scalar cache.
// Both fuse the residual (apply - rho) and an FP64 sum of |result|^2 per
// block (deterministic tree order) for the multigrid convergence test.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.h"
#include "stencil_box.h"
#include "stencil_tile.h"

namespace gpaw {

hipError_t box_launch_f64(const Stencil&, int, const void*, void*, const void*, const Geom&,
                          const Phases&, double, const TileShape&, double*, bool, hipStream_t);
hipError_t box_launch_c128(const Stencil&, int, const void*, void*, const void*, const Geom&,
                           const Phases&, double, const TileShape&, double*, bool, hipStream_t);
hipError_t box_launch_f32(const Stencil&, int, const void*, void*, const void*, const Geom&,
                          const Phases&, double, const TileShape&, double*, bool, hipStream_t);
hipError_t box_launch_c64(const Stencil&, int, const void*, void*, const void*, const Geom&,
                          const Phases&, double, const TileShape&, double*, bool, hipStream_t);

hipError_t tile_launch_f64(const Stencil&, int, const void*, void*, const void*, const Geom&,
                          const Phases&, double, int, const TileShape&, double*, bool, hipStream_t);
hipError_t tile_launch_c128(const Stencil&, int, const void*, void*, const void*, const Geom&,
                           const Phases&, double, int, const TileShape&, double*, bool, hipStream_t);
hipError_t tile_launch_f32(const Stencil&, int, const void*, void*, const void*, const Geom&,
                          const Phases&, double, int, const TileShape&, double*, bool, hipStream_t);
hipError_t tile_launch_c64(const Stencil&, int, const void*, void*, const void*, const Geom&,
                          const Phases&, double, int, const TileShape&, double*, bool, hipStream_t);

struct StarCoefs {
    double c[6 * GPAW_MAXR + 1];
};

template <class V>
__device__ __forceinline__ double norm2(V v);
template <> __device__ __forceinline__ double norm2<double>(double v) { return v * v; }
template <> __device__ __forceinline__ double norm2<float>(float v) { return (double)v * (double)v; }
template <> __device__ __forceinline__ double norm2<double2>(double2 v) { return v.x * v.x + v.y * v.y; }
template <> __device__ __forceinline__ double norm2<float2>(float2 v)
{
    return (double)v.x * (double)v.x + (double)v.y * (double)v.y;
}

// load with a precomputed wrap flag for one axis (0 inside, 1 high, 2 low, 3 zero)
template <class V>
__device__ __forceinline__ V flagged_load(const V* __restrict__ p, int flag, const Phases& ph, int axis)
{
    if (flag == 3) return vzero<V>();
    V v = *p;
    if (VT<V>::cplx && flag) v = pside(v, ph, axis, flag);
    return v;
}

template <class V, int R, int MODE, bool NORM, bool WRITE>
__global__ __launch_bounds__(256) void k_star(const V* __restrict__ in, V* __restrict__ out,
                                              const V* __restrict__ aux, Geom g, Phases ph,
                                              StarCoefs sc, double w, int q, int ntz, i64 ng,
                                              double* __restrict__ partial)
{
    typedef typename VT<V>::R RT;
    const int tz = blockIdx.x % ntz, ty = blockIdx.x / ntz;
    const int i2 = tz * blockDim.x + threadIdx.x;
    const int i1 = ty * blockDim.y + threadIdx.y;
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    const bool active = (i1 < n1) && (i2 < n2);
    double acc = 0.0;
    if (active) {
        const i64 plane = (i64)n1 * n2;
        const i64 boff = (i64)blockIdx.z * ng;
        const V* __restrict__ a = in + boff;
        const int i0s = blockIdx.y * q;
        const int i0e = min(i0s + q, n0);
        int zc[2 * R], zf[2 * R], yc[2 * R], yf[2 * R];
#pragma unroll
        for (int k = 1; k <= R; k++) {
            int j = i2 + k;
            zf[k - 1] = wrap_index(j, n2, g.per[2]);
            zc[k - 1] = j;
            j = i2 - k;
            zf[R + k - 1] = wrap_index(j, n2, g.per[2]);
            zc[R + k - 1] = j;
            j = i1 + k;
            yf[k - 1] = wrap_index(j, n1, g.per[1]);
            yc[k - 1] = j * n2;
            j = i1 - k;
            yf[R + k - 1] = wrap_index(j, n1, g.per[1]);
            yc[R + k - 1] = j * n2;
        }
        const int row = i1 * n2;
        V qv[2 * R + 1];
#pragma unroll
        for (int k = 0; k < 2 * R; k++) {
            int p = i0s - R + k;
            const int f = wrap_index(p, n0, g.per[0]);
            qv[k] = flagged_load(a + (f == 3 ? 0 : (i64)p * plane + row + i2), f, ph, 0);
        }
        for (int i0 = i0s; i0 < i0e; i0++) {
            {
                int p = i0 + R;
                const int f = wrap_index(p, n0, g.per[0]);
                qv[2 * R] = flagged_load(a + (f == 3 ? 0 : (i64)p * plane + row + i2), f, ph, 0);
            }
            const V* __restrict__ pl = a + (i64)i0 * plane;
            V x = vzero<V>();
            if (MODE != ST_JACOBI) x = vmac(x, qv[R], (RT)sc.c[0]);
#pragma unroll
            for (int k = 0; k < 2 * R; k++)   // axis 2: +1..+R then -1..-R
                x = vmac(x, flagged_load(pl + row + zc[k], zf[k], ph, 2), (RT)sc.c[1 + k]);
#pragma unroll
            for (int k = 0; k < 2 * R; k++)   // axis 1
                x = vmac(x, flagged_load(pl + yc[k] + i2, yf[k], ph, 1),
                         (RT)sc.c[1 + 2 * R + k]);
#pragma unroll
            for (int k = 1; k <= R; k++)      // axis 0: +k from the queue
                x = vmac(x, qv[R + k], (RT)sc.c[4 * R + k]);
#pragma unroll
            for (int k = 1; k <= R; k++)      // axis 0: -k
                x = vmac(x, qv[R - k], (RT)sc.c[5 * R + k]);
            const i64 idx = boff + (i64)i0 * plane + row + i2;
            const V r = st_epilogue<MODE>(x, qv[R],
                                          MODE == ST_APPLY ? vzero<V>()
                                                           : aux_value<MODE>(aux, boff, idx - boff),
                                          (RT)w, (RT)sc.c[0]);
            if (NORM) acc += norm2(r);
            if (WRITE) out[idx] = r;
#pragma unroll
            for (int k = 0; k < 2 * R; k++) qv[k] = qv[k + 1];
        }
    }
    if (NORM) {
        __shared__ double red[256];
        const double s = block_sum<256>(acc, red);
        if (threadIdx.x == 0 && threadIdx.y == 0)
            partial[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = s;
    }
}

template <class V, int MODE, bool NORM, bool WRITE>
__global__ __launch_bounds__(256) void k_generic(const V* __restrict__ in, V* __restrict__ out,
                                                 const V* __restrict__ aux, Geom g, Phases ph,
                                                 const double* __restrict__ coefs,
                                                 const int* __restrict__ off, int nc, double w,
                                                 i64 ng, double* __restrict__ partial)
{
    typedef typename VT<V>::R RT;
    const i64 t = (i64)blockIdx.x * 256 + threadIdx.x;
    double acc = 0.0;
    if (t < ng) {
        const int n1 = g.n[1], n2 = g.n[2];
        const int i2 = (int)(t % n2);
        const int i1 = (int)((t / n2) % n1);
        const int i0 = (int)(t / ((i64)n1 * n2));
        const i64 boff = (i64)blockIdx.y * ng;
        const V* __restrict__ a = in + boff;
        V x = vzero<V>();
        for (int c = (MODE == ST_JACOBI ? 1 : 0); c < nc; c++)
            x = vmac(x, ghost_fetch(a, i0 + off[3 * c], i1 + off[3 * c + 1], i2 + off[3 * c + 2],
                                    g, ph),
                     (RT)coefs[c]);
        const V r = st_epilogue<MODE>(x, MODE == ST_APPLY || MODE == ST_RESID ? vzero<V>() : a[t],
                                      MODE == ST_APPLY ? vzero<V>() : aux_value<MODE>(aux, boff, t), (RT)w,
                                      (RT)coefs[0]);
        if (NORM) acc = norm2(r);
        if (WRITE) out[boff + t] = r;
    }
    if (NORM) {
        __shared__ double red[256];
        const double s = block_sum<256>(acc, red);
        if (threadIdx.x == 0) partial[blockIdx.x + gridDim.x * blockIdx.y] = s;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

struct StarShape {
    dim3 grid, block;
    int q, ntz;
};

StarShape star_shape(const Geom& g, int nin)
{
    StarShape s;
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    // widest row segment that wastes at most ~10% of the lanes
    int bx = 8;
    for (int c : {64, 32, 16}) {
        const int tiles = (n2 + c - 1) / c;
        if (tiles * c - n2 <= (tiles * c) / 10 || c == 16) {
            bx = c;
            break;
        }
    }
    if (n2 <= 8) bx = 8;
    int by = 256 / bx;
    s.block = dim3(bx, by, 1);
    s.ntz = (n2 + bx - 1) / bx;
    const int nty = (n1 + by - 1) / by;
    const long base = (long)s.ntz * nty * nin;
    // enough blocks to fill 256 CUs several times; short i0 runs on small grids
    long chunks = (2048 + base - 1) / base;
    if (chunks > n0) chunks = n0;
    if (chunks < 1) chunks = 1;
    s.q = (int)((n0 + chunks - 1) / chunks);
    s.grid = dim3(s.ntz * nty, (n0 + s.q - 1) / s.q, nin);
    return s;
}

template <class V, int R, int MODE>
hipError_t star_launch(const Stencil& s, const void* in, void* out, const void* aux, const Geom& g,
                       const Phases& ph, int nin, double w, double* partial, bool write,
                       hipStream_t st)
{
    StarShape sh = star_shape(g, nin);
    StarCoefs sc;
    for (int c = 0; c < 6 * R + 1; c++) sc.c[c] = s.coefs[c];
    const i64 ng = (i64)g.n[0] * g.n[1] * g.n[2];
    const V* i = (const V*)in;
    V* o = (V*)out;
    const V* x = (const V*)aux;
    if constexpr (st_norm_ok<V, MODE>()) {
        if (partial) {
            if (write)
                launch_kernel((k_star<V, R, MODE, true, true>), sh.grid, sh.block, 0, st, i, o,
                                   x, g, ph, sc, w, sh.q, sh.ntz, ng, partial);
            else
                launch_kernel((k_star<V, R, MODE, true, false>), sh.grid, sh.block, 0, st, i,
                                   o, x, g, ph, sc, w, sh.q, sh.ntz, ng, partial);
            return launch_error();
        }
    }
    launch_kernel((k_star<V, R, MODE, false, true>), sh.grid, sh.block, 0, st, i, o, x, g, ph,
                       sc, w, sh.q, sh.ntz, ng, partial);
    return launch_error();
}

template <class V, int MODE>
hipError_t star_dispatch_r(const Stencil& s, const void* in, void* out, const void* aux,
                           const Geom& g, const Phases& ph, int nin, double w, double* partial,
                           bool write, hipStream_t st)
{
    switch (s.star_r) {
    case 1: return star_launch<V, 1, MODE>(s, in, out, aux, g, ph, nin, w, partial, write, st);
    case 2: return star_launch<V, 2, MODE>(s, in, out, aux, g, ph, nin, w, partial, write, st);
    case 3: return star_launch<V, 3, MODE>(s, in, out, aux, g, ph, nin, w, partial, write, st);
    case 4: return star_launch<V, 4, MODE>(s, in, out, aux, g, ph, nin, w, partial, write, st);
    case 5: return star_launch<V, 5, MODE>(s, in, out, aux, g, ph, nin, w, partial, write, st);
    case 6: return star_launch<V, 6, MODE>(s, in, out, aux, g, ph, nin, w, partial, write, st);
    }
    return hipErrorInvalidValue;
}

template <class V, int MODE>
hipError_t generic_launch(const Stencil& s, const void* in, void* out, const void* aux,
                          const Geom& g, const Phases& ph, int nin, double w, double* partial,
                          bool write, hipStream_t st)
{
    const i64 ng = (i64)g.n[0] * g.n[1] * g.n[2];
    dim3 grid((unsigned)((ng + 255) / 256), nin, 1);
    const V* i = (const V*)in;
    V* o = (V*)out;
    const V* x = (const V*)aux;
    if constexpr (st_norm_ok<V, MODE>()) {
        if (partial) {
            if (write)
                launch_kernel((k_generic<V, MODE, true, true>), grid, dim3(256), 0, st, i, o,
                                   x, g, ph, s.d_coefs, s.d_off, s.nc, w, ng, partial);
            else
                launch_kernel((k_generic<V, MODE, true, false>), grid, dim3(256), 0, st, i, o,
                                   x, g, ph, s.d_coefs, s.d_off, s.nc, w, ng, partial);
            return launch_error();
        }
    }
    launch_kernel((k_generic<V, MODE, false, true>), grid, dim3(256), 0, st, i, o, x, g, ph,
                       s.d_coefs, s.d_off, s.nc, w, ng, partial);
    return launch_error();
}

template <class V, int MODE>
hipError_t kind_launch(bool star, const Stencil& s, const void* in, void* out, const void* aux,
                       const Geom& g, const Phases& ph, int nin, double w, double* partial,
                       bool write, hipStream_t st)
{
    if constexpr (!st_mode_ok<V, MODE>()) {
        return hipErrorInvalidValue;
    } else {
        if (partial && !st_norm_ok<V, MODE>()) return hipErrorInvalidValue;
        if (star) return star_dispatch_r<V, MODE>(s, in, out, aux, g, ph, nin, w, partial, write, st);
        return generic_launch<V, MODE>(s, in, out, aux, g, ph, nin, w, partial, write, st);
    }
}

template <class V>
hipError_t dispatch_mode(bool star, const Stencil& s, int mode, const void* in, void* out,
                         const void* aux, const Geom& g, const Phases& ph, int nin, double w,
                         double* partial, bool write, hipStream_t st)
{
    switch (mode) {
    case ST_APPLY: return kind_launch<V, ST_APPLY>(star, s, in, out, aux, g, ph, nin, w, partial, write, st);
    case ST_RESID: return kind_launch<V, ST_RESID>(star, s, in, out, aux, g, ph, nin, w, partial, write, st);
    case ST_JACOBI: return kind_launch<V, ST_JACOBI>(star, s, in, out, aux, g, ph, nin, w, partial, write, st);
    case ST_PCUPD: return kind_launch<V, ST_PCUPD>(star, s, in, out, aux, g, ph, nin, w, partial, write, st);
    case ST_PCFIN: return kind_launch<V, ST_PCFIN>(star, s, in, out, aux, g, ph, nin, w, partial, write, st);
    case ST_HPSI: return kind_launch<V, ST_HPSI>(star, s, in, out, aux, g, ph, nin, w, partial, write, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace

// Kernel choice.  GPAW_AMD_STENCIL=tile|box|queue|generic forces one
// (timing experiments); the default is the one-shot box kernel for real
// data on grids that sit in the caches, the LDS tile kernel on other large
// grids, the register-queue kernel on other star stencils, the generic one
// otherwise.
enum { K_GENERIC = 0, K_QUEUE = 1, K_TILE = 2, K_BOX = 3 };

static int forced_kind()
{
    static int f = -2;
    if (f == -2) {
        const char* e = getenv("GPAW_AMD_STENCIL");
        f = !e ? -1
               : e[0] == 't' ? K_TILE
               : e[0] == 'b' ? K_BOX
               : e[0] == 'q' ? K_QUEUE
               : e[0] == 'g' ? K_GENERIC
                             : -1;
    }
    return f;
}

// Kernel family for launches without partial sums, set at run time
// (gpaw_stencil_kernel: tests and timing experiments); -1 = the default
// choice.  Launches with partial sums keep the default, whose block count
// sizes the callers' partial buffers.
static std::atomic<int> g_kind_override{-1};

// Tile shape forced for launches without partial sums (gpaw_tile_shape):
// threads, columns, planes; threads = 0: none.
static std::atomic<int> g_tile_nt{0}, g_tile_tz{32}, g_tile_q{1};

int tile_shape_set(int nt, int tz, int q)
{
    if (nt == 0) {
        g_tile_nt.store(0);
        return 0;
    }
    if ((nt != TILE_NT && nt != TILE_NT_MID && nt != TILE_NT_WIDE) || (tz != 32 && tz != 64) ||
        q < 1)
        return -1;
    g_tile_tz.store(tz);
    g_tile_q.store(q);
    g_tile_nt.store(nt);
    return 0;
}

int stencil_kernel_set(int kind)
{
    const int prev = g_kind_override.load();
    if (kind >= -1 && kind <= K_BOX) g_kind_override.store(kind);
    return prev;
}

static const char* kind_name(int k)
{
    return k == K_TILE ? "tile" : k == K_BOX ? "box" : k == K_QUEUE ? "queue" : "generic";
}

// Box kernel shapes (stencil_box.h): 32 x 8 columns, Q planes per block.
// Compiled for the (type, mode) pairs the tile kernel has.
bool box_mode_ok(int dtype, int mode)
{
    const bool cplx = dtype_complex(dtype), f64 = dtype == DT_F64 || dtype == DT_C128;
    switch (mode) {
    case ST_APPLY:
    case ST_RESID: return true;
    case ST_JACOBI: return !cplx;
    case ST_PCUPD:
    case ST_PCFIN:
    case ST_HPSI: return f64;
    }
    return false;
}

bool box_geometry(const Geom& g, int R, int nin, int Q, TileShape& sh)
{
    if (R < 1 || R > 4 || (Q != 2 && Q != 4 && Q != 8) || nin < 1) return false;
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    sh.nt = BOX_NT;
    sh.ntz = (n2 + BOX_TZ - 1) / BOX_TZ;
    sh.nty = (n1 + BOX_NTY - 1) / BOX_NTY;
    // every coordinate a block touches lies within one period on either side
    // of the grid (the last chunk's column reaches R planes past its Q)
    if (sh.ntz * BOX_TZ + R > 2 * n2 || sh.nty * BOX_NTY + R > 2 * n1 || n0 < Q + R ||
        n1 < R || n2 < R)
        return false;
    if ((long)n1 * n2 >= PLANE_REF_MAX) return false;
    sh.q = Q;
    sh.nchunk = (n0 + Q - 1) / Q;
    const long nb = (long)sh.ntz * sh.nty * nin * sh.nchunk;
    if (nb > 0x7fffffffL) return false;
    sh.nblocks = (int)nb;
    return true;
}

// Grids of up to this many points (all bands) take the box kernel by
// default: there the tile kernel's per-plane round trips dominate.
constexpr long BOX_MAX_POINTS = 1L << 23;

int box_default_q(const Geom& g, int R, int nin)
{
    (void)g;
    (void)R;
    (void)nin;
    return 4;
}

static bool box_default(const Stencil& s, const Geom& g, int dtype, int mode, int nin,
                        TileShape& sh)
{
    // real data: the multigrid levels and their residual / norm passes
    if (dtype_complex(dtype) || !box_mode_ok(dtype, mode) || s.star_r < 1 || s.star_r > 4)
        return false;
    if ((long)g.n[0] * g.n[1] * g.n[2] * nin > BOX_MAX_POINTS) return false;
    return box_geometry(g, s.star_r, nin, box_default_q(g, s.star_r, nin), sh);
}

hipError_t box_launch(const Stencil& s, int dtype, int mode, const void* in, void* out,
                      const void* aux, const Geom& g, const Phases& ph, double w,
                      const TileShape& sh, double* partial, bool write, hipStream_t st)
{
    if (!box_mode_ok(dtype, mode)) return hipErrorInvalidValue;
    switch (dtype) {
    case DT_F64: return box_launch_f64(s, mode, in, out, aux, g, ph, w, sh, partial, write, st);
    case DT_C128: return box_launch_c128(s, mode, in, out, aux, g, ph, w, sh, partial, write, st);
    case DT_F32: return box_launch_f32(s, mode, in, out, aux, g, ph, w, sh, partial, write, st);
    case DT_C64: return box_launch_c64(s, mode, in, out, aux, g, ph, w, sh, partial, write, st);
    }
    return hipErrorInvalidValue;
}

bool jacobi_two_from_zero_ok(const Stencil& s, int dtype, const Geom& g)
{
    TileShape sh;
    return s.star_r >= 1 && s.star_r <= 3 && (dtype == DT_F64 || dtype == DT_F32) &&
           box_geometry(g, s.star_r, 1, box_default_q(g, s.star_r, 1), sh);
}

hipError_t launch_jacobi_two_from_zero(const Stencil& s, int dtype, const void* rho, void* out,
                                       const Geom& g, double w0, double w1, hipStream_t st)
{
    if (!jacobi_two_from_zero_ok(s, dtype, g)) return hipErrorInvalidValue;
    TileShape sh;
    box_geometry(g, s.star_r, 1, box_default_q(g, s.star_r, 1), sh);
    return dtype == DT_F64 ? box_fz_f64(s, rho, out, g, w0, w1, sh, st)
                           : box_fz_f32(s, rho, out, g, w0, w1, sh, st);
}

int tile_tz(int)
{
    // 32 columns (8 waves of rows per block): measured faster than 64 on
    // every grid tried (64^3, 95^3; float64 and complex128), and it wastes
    // fewer lanes on edges that are not multiples of 64
    return 32;
}

bool tile_geometry(const Geom& g, int R, int esize, int nin, int TZ, int q, TileShape& sh,
                   int nt)
{
    if (R < 1 || R > 4 || (TZ != 32 && TZ != 64) || q < 1 || nin < 1) return false;
    if (nt != TILE_NT && nt != TILE_NT_WIDE && nt != TILE_NT_MID) return false;
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    const int TY = (nt / TZ) * tile_rows(R, esize, nt);
    sh.nt = nt;
    sh.ntz = (n2 + TZ - 1) / TZ;
    sh.nty = (n1 + TY - 1) / TY;
    // every coordinate a tile touches (tile + halo) must lie within one
    // period on either side of the grid: ghost indices wrap only once
    if (sh.ntz * TZ + R > 2 * n2 || sh.nty * TY + R > 2 * n1 || n0 < R || n1 < R || n2 < R)
        return false;
    if ((long)n1 * n2 >= PLANE_REF_MAX) return false;   // packed in-plane offsets
    sh.q = q < n0 ? q : n0;
    sh.nchunk = (n0 + sh.q - 1) / sh.q;
    const long nb = (long)sh.ntz * sh.nty * nin * sh.nchunk;
    if (nb > 0x7fffffffL) return false;
    sh.nblocks = (int)nb;
    return true;
}

int tile_default_q(const Geom& g, int R, int esize, int nin, int TZ, int nt)
{
    // enough tiles to fill the chip twice at 8 blocks per CU, runs of >= 8
    // planes (each run re-reads 2R planes of queue)
    const int n0 = g.n[0];
    const int TY = (nt / TZ) * tile_rows(R, esize, nt);
    const long ntile = (long)((g.n[2] + TZ - 1) / TZ) * ((g.n[1] + TY - 1) / TY) * nin;
    long nch = (4096 * TILE_NT / nt + ntile - 1) / ntile;
    const long maxch = n0 / 8 > 1 ? n0 / 8 : 1;
    if (nch > maxch) nch = maxch;
    if (nch < 1) nch = 1;
    return (int)((n0 + nch - 1) / nch);
}

bool tile_shape(const Geom& g, int R, int esize, int nin, int& TZ, TileShape& sh, int nt)
{
    if (R < 1 || R > 4) return false;
    TZ = tile_tz(g.n[2]);
    if (!tile_geometry(g, R, esize, nin, TZ, tile_default_q(g, R, esize, nin, TZ, nt), sh, nt))
        return false;
    return (long)sh.nblocks * nt >= 256L * TILE_NT;
}

// `dtype` and `mode` decide whether the box kernel applies; partial-sum
// layouts are asked for with DT_F64 / ST_RESID (norms are taken of real
// data only, and float and double grids get the same box and tile shapes)
static int stencil_kind(const Stencil& s, const Geom& g, int dtype, int mode, int nin, int& TZ,
                        TileShape& sh, int nt = TILE_NT)
{
    const int f = forced_kind();
    const int esize = dtype_size(dtype);
    if (f == K_GENERIC && s.d_coefs) return K_GENERIC;
    if (f == K_QUEUE && s.star_r > 0) return K_QUEUE;
    if (f != K_QUEUE && f != K_TILE && s.star_r > 0 && box_mode_ok(dtype, mode)) {
        TileShape bs;
        if (f == K_BOX ? box_geometry(g, s.star_r, nin, box_default_q(g, s.star_r, nin), bs)
                       : box_default(s, g, dtype, mode, nin, bs)) {
            TZ = BOX_TZ;
            sh = bs;
            return K_BOX;
        }
    }
    const bool tile_ok = s.star_r > 0 && tile_shape(g, s.star_r, esize, nin, TZ, sh, nt);
    if (tile_ok && f != K_QUEUE) return K_TILE;
    return s.star_r > 0 ? K_QUEUE : K_GENERIC;
}

int stencil_partials(const Stencil& s, const Geom& g, int nin)
{
    // norms are taken of real data only, whose tiles all have the same
    // shape (4- and 8-byte elements)
    int TZ;
    TileShape sh;
    const int kind = stencil_kind(s, g, DT_F64, ST_RESID, nin, TZ, sh);
    if (kind == K_TILE || kind == K_BOX) return sh.nblocks;
    if (kind == K_QUEUE) {
        StarShape q = star_shape(g, nin);
        return q.grid.x * q.grid.y * q.grid.z;
    }
    const i64 ng = (i64)g.n[0] * g.n[1] * g.n[2];
    return (int)((ng + 255) / 256) * nin;
}

// One launch configuration: the kernel kind and, for tiles, the shape.
struct LaunchCfg {
    int kind;
    int TZ;
    TileShape sh;
};

static hipError_t launch_cfg(const LaunchCfg& c, const Stencil& s, int dtype, int mode,
                             const void* in, void* out, const void* aux, const Geom& g,
                             const Phases& ph, int nin, double w, double* partial, bool write,
                             hipStream_t st)
{
    if (c.kind == K_BOX)
        return box_launch(s, dtype, mode, in, out, aux, g, ph, w, c.sh, partial, write, st);
    if (c.kind == K_TILE) {
        switch (dtype) {
        case DT_F64: return tile_launch_f64(s, mode, in, out, aux, g, ph, w, c.TZ, c.sh, partial, write, st);
        case DT_C128: return tile_launch_c128(s, mode, in, out, aux, g, ph, w, c.TZ, c.sh, partial, write, st);
        case DT_F32: return tile_launch_f32(s, mode, in, out, aux, g, ph, w, c.TZ, c.sh, partial, write, st);
        case DT_C64: return tile_launch_c64(s, mode, in, out, aux, g, ph, w, c.TZ, c.sh, partial, write, st);
        }
        return hipErrorInvalidValue;
    }
    const bool star = c.kind == K_QUEUE;
    switch (dtype) {
    case DT_F64: return dispatch_mode<double>(star, s, mode, in, out, aux, g, ph, nin, w, partial, write, st);
    case DT_C128: return dispatch_mode<double2>(star, s, mode, in, out, aux, g, ph, nin, w, partial, write, st);
    case DT_F32: return dispatch_mode<float>(star, s, mode, in, out, aux, g, ph, nin, w, partial, write, st);
    case DT_C64: return dispatch_mode<float2>(star, s, mode, in, out, aux, g, ph, nin, w, partial, write, st);
    }
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Autotuner.  For a star stencil of range 1..4 the launch shape -- the
// register-queue kernel, or the tile kernel with TZ = 32 / 64 columns and
// q = 1 .. n0 planes per tile -- is chosen by timing every candidate on the
// caller's own launch the first time a (type, mode, range, grid, bands)
// combination is seen.  Every shape computes identical bits and input and
// output never alias, so running the candidates repeats the same result.
// Off by default (heuristic shapes); gpaw_autotune(1) or GPAW_AMD_AUTOTUNE=1
// turns it on.  Launches that produce norm partials (whose count is fixed
// by the heuristic shape) and launches recorded into a graph are not tuned;
// the latter use an existing entry.
// ---------------------------------------------------------------------------
namespace {

struct TuneKey {
    int dtype, mode, R, n0, n1, n2, per, nin;
    bool operator<(const TuneKey& o) const
    {
        const int a[8] = {dtype, mode, R, n0, n1, n2, per, nin};
        const int b[8] = {o.dtype, o.mode, o.R, o.n0, o.n1, o.n2, o.per, o.nin};
        return std::lexicographical_compare(a, a + 8, b, b + 8);
    }
};

struct TuneVal {
    LaunchCfg cfg;
    float us;
    int candidates;
};

std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune;
std::atomic<int> g_tune_on{-1};

bool tune_enabled()
{
    int v = g_tune_on.load();
    if (v < 0) {
        const char* e = getenv("GPAW_AMD_AUTOTUNE");
        v = (e && atoi(e) > 0) ? 1 : 0;
        g_tune_on.store(v);
    }
    return v > 0;
}

// Mean time of one candidate's launch, replayed back to back in a hipGraph
// as the multigrid plans run them (eager launches add host submission time,
// which varies by more than the differences being measured).  One untimed
// eager launch first checks the candidate and gives a rough duration, from
// which the graph gets enough launches for ~300 us.
float time_cfg(const LaunchCfg& c, const Stencil& s, int dtype, int mode, const void* in,
               void* out, const void* aux, const Geom& g, const Phases& ph, int nin, double w,
               bool write, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    hipEventRecord(e0, st);
    if (launch_cfg(c, s, dtype, mode, in, out, aux, g, ph, nin, w, nullptr, write, st) != hipSuccess)
        return -1.0f;
    hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess || hipGetLastError() != hipSuccess) return -1.0f;
    float ms = 0.0f;
    hipEventElapsedTime(&ms, e0, e1);
    int reps = ms > 0.0f ? (int)(0.3f / ms) : 32;
    reps = reps < 2 ? 2 : reps > 32 ? 32 : reps;
    hipGraphExec_t ge = nullptr;
    if (build_graph(&ge, [&](hipStream_t) {
            hipError_t e = hipSuccess;
            for (int r = 0; r < reps && e == hipSuccess; r++)
                e = launch_cfg(c, s, dtype, mode, in, out, aux, g, ph, nin, w, nullptr, write,
                               nullptr);
            return e;
        }) != hipSuccess)
        return -1.0f;
    float best = -1.0f;
    for (int pass = 0; pass < 4; pass++) {   // first replay untimed
        hipEventRecord(e0, st);
        hipError_t e = hipGraphLaunch(ge, st);
        hipEventRecord(e1, st);
        if (e != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            best = -1.0f;
            break;
        }
        hipEventElapsedTime(&ms, e0, e1);
        const float us = 1e3f * ms / reps;
        if (pass > 0 && (best < 0 || us < best)) best = us;
    }
    hipGraphExecDestroy(ge);
    return best;
}

TuneVal tune(const TuneKey& key, const LaunchCfg& heuristic, const Stencil& s, int dtype,
             int mode, const void* in, void* out, const void* aux, const Geom& g,
             const Phases& ph, int nin, double w, bool write, hipStream_t st)
{
    std::vector<LaunchCfg> cand;
    LaunchCfg qc{K_QUEUE, 64, TileShape{}};
    cand.push_back(qc);
    const bool wide = dtype_complex(dtype) && s.star_r <= 3;
    for (int nt : {TILE_NT, TILE_NT_MID, TILE_NT_WIDE}) {
        if (nt != TILE_NT && !wide) continue;
        for (int TZ : {64, 32})
            for (int q : {1, 2, 4, 8, 16, 32, 64, 1 << 20}) {
                LaunchCfg c{K_TILE, TZ, TileShape{}};
                if (!tile_geometry(g, s.star_r, dtype_size(dtype), nin, TZ, q, c.sh, nt)) continue;
                bool dup = false;
                for (const LaunchCfg& o : cand)
                    dup |= o.kind == K_TILE && o.TZ == TZ && o.sh.q == c.sh.q && o.sh.nt == nt;
                if (!dup) cand.push_back(c);
            }
    }
    if (box_mode_ok(dtype, mode))
        for (int q : {2, 4, 8}) {
            LaunchCfg c{K_BOX, BOX_TZ, TileShape{}};
            if (box_geometry(g, s.star_r, nin, q, c.sh)) cand.push_back(c);
        }
    TuneVal best{heuristic, -1.0f, 0};
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return best;
    if (hipEventCreate(&e1) != hipSuccess) {
        hipEventDestroy(e0);
        return best;
    }
    // GPAW_AMD_TUNE_LOG=1: every candidate's time on stderr
    static const bool log = [] {
        const char* e = getenv("GPAW_AMD_TUNE_LOG");
        return e && atoi(e) > 0;
    }();
    for (const LaunchCfg& c : cand) {
        const float us = time_cfg(c, s, dtype, mode, in, out, aux, g, ph, nin, w, write, st, e0, e1);
        if (log)
            fprintf(stderr,
                    "tune dtype=%d mode=%d R=%d n=%dx%dx%d nin=%d: %s TZ=%d q=%d threads=%d "
                    "%.2f us\n",
                    dtype, mode, s.star_r, g.n[0], g.n[1], g.n[2], nin, kind_name(c.kind),
                    c.TZ, c.kind >= K_TILE ? c.sh.q : 0, c.kind >= K_TILE ? c.sh.nt : 256, us);
        if (us < 0) continue;
        best.candidates++;
        if (best.us < 0 || us < best.us) {
            best.cfg = c;
            best.us = us;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    (void)key;
    return best;
}

}  // namespace

int autotune_set(int on)
{
    const int prev = tune_enabled() ? 1 : 0;
    if (on >= 0) g_tune_on.store(on > 0 ? 1 : 0);
    return prev;
}

long autotune_report(char* buf, long len)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::string r;
    char line[256];
    for (const auto& kv : g_tune) {
        const TuneKey& k = kv.first;
        const TuneVal& v = kv.second;
        snprintf(line, sizeof line,
                 "dtype=%d mode=%d R=%d n=%dx%dx%d periodic=%d nin=%d -> %s TZ=%d q=%d "
                 "threads=%d %.2f us (%d candidates)\n",
                 k.dtype, k.mode, k.R, k.n0, k.n1, k.n2, k.per, k.nin, kind_name(v.cfg.kind),
                 v.cfg.kind >= K_TILE ? v.cfg.TZ : 0, v.cfg.kind >= K_TILE ? v.cfg.sh.q : 0,
                 v.cfg.kind >= K_TILE ? v.cfg.sh.nt : 256, v.us, v.candidates);
        r += line;
    }
    if (buf && len > 0) {
        const long n = (long)r.size() < len - 1 ? (long)r.size() : len - 1;
        memcpy(buf, r.data(), n);
        buf[n] = 0;
    }
    return (long)r.size();
}

hipError_t launch_stencil(const Stencil& s, int dtype, int mode, const void* in, void* out,
                          const void* aux, const Geom& g, const Phases& ph, int nin, double w,
                          double* partial, int* npartial, bool write, hipStream_t st)
{
    LaunchCfg c{K_GENERIC, 64, TileShape{}};
    // complex applies of range <= 3 use 1024-thread tiles
    const int nt = (dtype_complex(dtype) && (mode == ST_APPLY || mode == ST_HPSI) && s.star_r <= 3)
                       ? TILE_NT_WIDE
                       : TILE_NT;
    c.kind = stencil_kind(s, g, dtype, mode, nin, c.TZ, c.sh, nt);
    if (npartial) *npartial = stencil_partials(s, g, nin);
    // 4-byte real data on grids of 2^19 points or more: boxes of 8 planes
    // (measured 6.0 / 9.2 / 14.5 us against 6.4 / 10.6 / 17.9 us for 4 planes
    // on 95^3 / 127^3 / 159^3); launches with partial sums keep the shared
    // layout above, whose block count sizes the callers' buffers
    if (!partial && c.kind == K_BOX && dtype == DT_F32 &&
        (long)g.n[0] * g.n[1] * g.n[2] * nin >= (1L << 19)) {
        TileShape b8;
        if (box_geometry(g, s.star_r, nin, 8, b8)) c.sh = b8;
    }
    const int ov = g_kind_override.load();
    if (!partial && ov >= 0 && ov != c.kind) {
        // the requested family when it can run this launch, else the default
        LaunchCfg o{ov, 64, TileShape{}};
        bool ok = false;
        if (ov == K_BOX) {
            ok = box_mode_ok(dtype, mode) && s.star_r >= 1 && s.star_r <= 4 &&
                 box_geometry(g, s.star_r, nin, box_default_q(g, s.star_r, nin), o.sh);
            o.TZ = BOX_TZ;
        } else if (ov == K_TILE) {
            ok = s.star_r >= 1 && tile_shape(g, s.star_r, dtype_size(dtype), nin, o.TZ, o.sh, nt);
            const int fnt = g_tile_nt.load();
            // a forced shape where the kernel is compiled for it (wide blocks:
            // complex data of range <= 3)
            if (ok && fnt && (fnt == TILE_NT || (dtype_complex(dtype) && s.star_r <= 3))) {
                TileShape fs;
                if (tile_geometry(g, s.star_r, dtype_size(dtype), nin, g_tile_tz.load(),
                                  g_tile_q.load(), fs, fnt)) {
                    o.TZ = g_tile_tz.load();
                    o.sh = fs;
                }
            }
        } else if (ov == K_QUEUE) {
            ok = s.star_r > 0;
        } else {
            ok = s.d_coefs != nullptr;
        }
        if (ok) return launch_cfg(o, s, dtype, mode, in, out, aux, g, ph, nin, w, nullptr, write, st);
    }
    if (!partial && s.star_r >= 1 && s.star_r <= 4 && forced_kind() < 0 && tune_enabled()) {
        const TuneKey key{dtype, mode, s.star_r, g.n[0], g.n[1], g.n[2],
                          g.per[0] | g.per[1] << 1 | g.per[2] << 2, nin};
        bool found = false;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tune.find(key);
            if (it != g_tune.end()) {
                c = it->second.cfg;
                found = true;
            }
        }
        if (!found && !graph_recorder() && in != out) {
            const TuneVal v = tune(key, c, s, dtype, mode, in, out, aux, g, ph, nin, w, write, st);
            if (v.us >= 0) {
                std::lock_guard<std::mutex> lk(g_tune_mu);
                g_tune[key] = v;
                c = v.cfg;
            }
        }
    }
    // in . out partials of an apply and input-residual partials of a Jacobi
    // sweep come from the tile and box kernels only
    if (partial && (mode == ST_APPLY || mode == ST_JACOBI) && c.kind != K_TILE && c.kind != K_BOX)
        return hipErrorInvalidValue;
    return launch_cfg(c, s, dtype, mode, in, out, aux, g, ph, nin, w, partial, write, st);
}

bool stencil_dot_ok(const Stencil& s, int dtype, const Geom& g, int nin)
{
    int TZ;
    TileShape sh;
    const int k = stencil_kind(s, g, dtype, ST_RESID, nin, TZ, sh);
    return (dtype == DT_F64 || dtype == DT_F32) && (k == K_TILE || k == K_BOX);
}

}  // namespace gpaw
