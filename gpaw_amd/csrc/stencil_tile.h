// LDS-tiled 2.5-D star-stencil kernel for large grids (gfx950).
//
// A block of 256 threads (512 or 1024 for complex data of range <= 3) owns
// a TY x TZ tile of the (i1, i2) plane and sweeps a run of i0 planes.  Each
// thread owns ROWS consecutive i1 rows of one i2 column:
//   * axis-0 neighbours come from a register queue of 2R+1 planes per row
//     (every input value is read from HBM once per sweep);
//   * the current plane, with an R-wide halo in i1 and i2, is staged in LDS
//     (double-buffered, one barrier per plane);
//   * axis-1 neighbours use a register window of ROWS + 2R values per
//     column, so only the 2R window ends are LDS reads;
//   * the next plane's queue values and halo cells are loaded before the
//     barrier, overlapping their latency with the current plane's work
//     (distances of two and three planes were measured no faster).
// Waves are TZ lanes wide along the contiguous axis, so every HBM access is
// a full coalesced row segment.  Tiles are ordered so that one XCD (one L2)
// processes a contiguous range of tiles: the halo rows a block re-reads
// were loaded by its neighbour on the same XCD.
//
// The sum runs in the reference's coefficient order (centre, +-k along
// axis 2, axis 1, axis 0: c/bmgs/fd.c:41-44 with the NewGUCLaplace order),
// so results are bit-identical to the reference C.
#pragma once
#include "kernels.h"

namespace gpaw {

// threads per tile block: 256, or 1024 for complex applies of range <= 3
// (taller tiles, fewer halo rows: measured 7% faster on complex128 bands,
// slower on float64); 512 and 1024 are tuner candidates for every complex
// mode of range <= 3
constexpr int TILE_NT = 256;
constexpr int TILE_NT_WIDE = 1024;
constexpr int TILE_NT_MID = 512;
// planes by which halo loads lead their use beyond the one-plane prefetch
constexpr int tile_halo_lead(int R, int, int nt) { return nt > TILE_NT ? R : 0; }

struct TileCoefs {
    double c[6 * GPAW_MAXR + 1];
};

struct TileShape {
    int ntz, nty, nchunk, q;   // tiles along i2, i1; chunks along i0 of q planes
    int nblocks;
    int nt = TILE_NT;          // threads per block
};

// Output store of the tile sweep.  NT: non-temporal (streaming) store,
// for outputs that are not read again soon (large complex applies).
template <bool NT, class V>
__device__ __forceinline__ void tile_store(V* p, V v)
{
    if constexpr (NT && VT<V>::cplx) {
        typedef typename VT<V>::R RT;
        typedef RT RT2 __attribute__((ext_vector_type(2)));
        __builtin_nontemporal_store(__builtin_bit_cast(RT2, v), reinterpret_cast<RT2*>(p));
    } else {
        *p = v;
    }
}

template <class V>
__device__ __forceinline__ double tnorm2(V v);
// in . result for real data (norm passes of ST_APPLY)
__device__ __forceinline__ double tdot(double a, double b) { return a * b; }
__device__ __forceinline__ double tdot(float a, float b) { return (double)a * (double)b; }
template <class V>
__device__ __forceinline__ double tdot(V, V) { return 0.0; }
template <> __device__ __forceinline__ double tnorm2<double>(double v) { return v * v; }
template <> __device__ __forceinline__ double tnorm2<float>(float v) { return (double)v * (double)v; }
template <> __device__ __forceinline__ double tnorm2<double2>(double2 v) { return v.x * v.x + v.y * v.y; }
template <> __device__ __forceinline__ double tnorm2<float2>(float2 v)
{
    return (double)v.x * (double)v.x + (double)v.y * (double)v.y;
}

template <class V, int R, int TZ, int ROWS, int NT = TILE_NT>
struct TileGeom {
    static constexpr int NTY = NT / TZ;
    static constexpr int TY = NTY * ROWS;
    static constexpr int LW = TZ + 2 * R;              // LDS row length
    static constexpr int LH = TY + 2 * R;              // LDS rows
    static constexpr int NHALO = 2 * R * (TZ + TY);    // halo cells per plane
    static constexpr int NH = (NHALO + NT - 1) / NT;   // halo cells per thread
    static constexpr int LDS = 2 * LH * LW;            // two plane buffers
};

// One tile of one band: columns z0 .. z0+TZ-1, rows y0 .. y0+TY-1, planes
// i0s .. i0e-1.  a / o / sa point at the band's input, output and aux
// arrays; lds holds TileGeom::LDS elements.  Every thread of the block
// calls it with the same tile.  Returns the thread's sum of |result|^2
// (NORM).
//
// Loads for the next plane are issued unconditionally from precomputed,
// always-valid addresses and only corrected (zero / Bloch phase) when they
// are consumed one plane later: no branch or phase lookup sits between a
// load and the barrier, so the load latency overlaps a whole plane of work.
template <class V, int R, int TZ, int ROWS, int MODE, bool NORM, bool WRITE, int NT = TILE_NT>
__device__ __forceinline__ double star_tile_sweep(const V* __restrict__ a, V* __restrict__ o,
                                                  const V* __restrict__ sa, const Geom g,
                                                  const Phases ph, const TileCoefs sc, double w,
                                                  int z0, int y0, int i0s, int i0e, V* lds)
{
    typedef typename VT<V>::R RT;
    typedef TileGeom<V, R, TZ, ROWS, NT> TG;
    constexpr int TY = TG::TY, LW = TG::LW, LH = TG::LH, NHALO = TG::NHALO, NH = TG::NH;
    // streaming stores for the wide complex applies (many bands: the output
    // is not re-read before it leaves the caches)
    constexpr bool NTS = VT<V>::cplx && MODE == ST_APPLY && NT == TILE_NT_WIDE;

    const int tx = threadIdx.x, tyi = threadIdx.y;
    const int tid = tx + TZ * tyi;
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    const i64 plane = (i64)n1 * n2;
    const int i2 = z0 + tx;

    // owned rows: output flag and in-plane reference (out-of-grid rows hold
    // the ghost values their in-grid neighbours read)
    bool inr[ROWS];
    PlaneRef qr[ROWS];
#pragma unroll
    for (int j = 0; j < ROWS; j++) {
        const int i1 = y0 + tyi * ROWS + j;
        inr[j] = (i1 < n1) && (i2 < n2);
        qr[j] = plane_ref(i1, i2, g);
    }
    // halo cells of this thread: LDS offset (-1: none) and in-plane reference
    int hoff[NH];
    PlaneRef hr[NH];
#pragma unroll
    for (int k = 0; k < NH; k++) {
        const int h = tid + NT * k;
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
        // threads without a halo cell load offset 0 of the plane (discarded)
        hr[k] = (h < NHALO) ? plane_ref(y0 + lr - R, z0 + lc - R, g) : 0;
    }
    // plane p: wrap flag of axis 0 (uniform) and the plane's base offset
    auto pflag = [&](int p, i64& base) -> int {
        const int f = wrap_index(p, n0, g.per[0]);
        base = f == 3 ? 0 : (i64)p * plane;
        return f;
    };

    // planes i0-R .. i0+R of the owned rows; the queue plane i0 + 1 + R is
    // loaded raw during plane i0 (flag qf).  The halo cells of a plane are
    // loaded together with that plane's queue rows (HA = R planes before
    // they are needed; slot p of hq holds plane i0 + p): the rows a
    // neighbouring tile loads for its queue are then still in the L2 when
    // they are re-read as halo.
    constexpr int HA = tile_halo_lead(R, (int)sizeof(V), NT);
    V q[ROWS][2 * R + 1];
    V qa[ROWS], hq[HA + 1][NH], sn[ROWS];
    int qf, hf[HA + 1];
#pragma unroll
    for (int k = 0; k < 2 * R; k++) {
        i64 base;
        const int f = pflag(i0s - R + k, base);
#pragma unroll
        for (int j = 0; j < ROWS; j++) q[j][k] = ghost_fix(a[base + pref_off(qr[j])], f, qr[j], ph);
    }
    {
        i64 base;
        qf = pflag(i0s + R, base);
#pragma unroll
        for (int j = 0; j < ROWS; j++) qa[j] = a[base + pref_off(qr[j])];
#pragma unroll
        for (int p = 0; p <= HA; p++) {
            hf[p] = pflag(i0s + p, base);
#pragma unroll
            for (int k = 0; k < NH; k++) hq[p][k] = a[base + pref_off(hr[k])];
        }
        if (MODE != ST_APPLY) {
#pragma unroll
            for (int j = 0; j < ROWS; j++)
                sn[j] = aux_value<MODE>(sa, 0, (i64)i0s * plane + (inr[j] ? pref_off(qr[j]) : 0));
        }
    }

    // Each plane step first consumes the loads issued one step earlier, then
    // issues the store of the previous plane's result and the next plane's
    // loads together, then computes.  Loads and stores share one counter
    // that does not retire them in order, so a wait for a load also waits
    // for every store issued before it: storing one step late keeps the
    // store out of the next wait's critical path.
    double acc = 0.0;
    V rp[ROWS];
    for (int i0 = i0s; i0 < i0e; i0++) {
        V* B = lds + ((i0 - i0s) & 1) * (LH * LW);
        // wait for all of the step's loads here, on every path: a load
        // consumed only inside a branch would make the next load into the
        // same registers wait again after the new loads were issued
#pragma unroll
        for (int j = 0; j < ROWS; j++) {
            vtouch(qa[j]);
            if (MODE != ST_APPLY) vtouch(sn[j]);
        }
#pragma unroll
        for (int p = 0; p <= HA; p++)
#pragma unroll
            for (int k = 0; k < NH; k++) vtouch(hq[p][k]);
        V s[ROWS];
#pragma unroll
        for (int j = 0; j < ROWS; j++) {
            q[j][2 * R] = ghost_fix(qa[j], qf, qr[j], ph);
            B[(R + tyi * ROWS + j) * LW + R + tx] = q[j][R];
            if (MODE != ST_APPLY) s[j] = sn[j];
        }
#pragma unroll
        for (int k = 0; k < NH; k++)
            if (hoff[k] >= 0) B[hoff[k]] = ghost_fix(hq[0][k], hf[0], hr[k], ph);
#pragma unroll
        for (int p = 0; p < HA; p++) {
            hf[p] = hf[p + 1];
#pragma unroll
            for (int k = 0; k < NH; k++) hq[p][k] = hq[p + 1][k];
        }
        if (WRITE && i0 > i0s) {
#pragma unroll
            for (int j = 0; j < ROWS; j++)
                if (inr[j]) tile_store<NTS>(o + (i64)(i0 - 1) * plane + pref_off(qr[j]), rp[j]);
        }
        if (i0 + 1 < i0e) {
            i64 base;
            qf = pflag(i0 + 1 + R, base);
#pragma unroll
            for (int j = 0; j < ROWS; j++) qa[j] = a[base + pref_off(qr[j])];
            if (i0 + 1 + HA < i0e) {
                if (HA) {   // same plane as the queue rows
                    hf[HA] = qf;
                } else {
                    hf[HA] = pflag(i0 + 1, base);
                }
#pragma unroll
                for (int k = 0; k < NH; k++) hq[HA][k] = a[base + pref_off(hr[k])];
            }
            if (MODE != ST_APPLY) {
#pragma unroll
                for (int j = 0; j < ROWS; j++)
                    sn[j] = aux_value<MODE>(sa, 0, (i64)(i0 + 1) * plane + (inr[j] ? pref_off(qr[j]) : 0));
            }
        }
        __syncthreads();
        V win[ROWS + 2 * R];
#pragma unroll
        for (int k = 0; k < R; k++) {
            win[k] = B[(tyi * ROWS + k) * LW + R + tx];
            win[ROWS + R + k] = B[(tyi * ROWS + ROWS + R + k) * LW + R + tx];
        }
#pragma unroll
        for (int j = 0; j < ROWS; j++) win[R + j] = q[j][R];
#pragma unroll
        for (int j = 0; j < ROWS; j++) {
            const V* row = B + (R + tyi * ROWS + j) * LW + R + tx;
            // RIN (Jacobi with NORM): also the residual of the input, summed
            // as the residual pass sums it (centre term first)
            constexpr bool RIN = NORM && MODE == ST_JACOBI;
            V x = vzero<V>(), xr = vzero<V>();
            if (MODE != ST_JACOBI) x = vmac(x, q[j][R], (RT)sc.c[0]);
            if (RIN) xr = vmac(xr, q[j][R], (RT)sc.c[0]);
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
                x = vmac(x, win[R + j + k], (RT)sc.c[2 * R + k]);
                if (RIN) xr = vmac(xr, win[R + j + k], (RT)sc.c[2 * R + k]);
            }
#pragma unroll
            for (int k = 1; k <= R; k++) {
                x = vmac(x, win[R + j - k], (RT)sc.c[3 * R + k]);
                if (RIN) xr = vmac(xr, win[R + j - k], (RT)sc.c[3 * R + k]);
            }
#pragma unroll
            for (int k = 1; k <= R; k++) {
                x = vmac(x, q[j][R + k], (RT)sc.c[4 * R + k]);
                if (RIN) xr = vmac(xr, q[j][R + k], (RT)sc.c[4 * R + k]);
            }
#pragma unroll
            for (int k = 1; k <= R; k++) {
                x = vmac(x, q[j][R - k], (RT)sc.c[5 * R + k]);
                if (RIN) xr = vmac(xr, q[j][R - k], (RT)sc.c[5 * R + k]);
            }
            rp[j] = st_epilogue<MODE>(x, q[j][R], MODE == ST_APPLY ? vzero<V>() : s[j], (RT)w,
                                      (RT)sc.c[0]);
            // NORM: |result|^2; for ST_APPLY in . result (a CG step's p.Lp);
            // for ST_JACOBI |L in - rho|^2 (the previous V-cycle's error)
            if (NORM && inr[j]) {
                if (MODE == ST_APPLY)
                    acc += tdot(q[j][R], rp[j]);
                else if (RIN)
                    acc += tnorm2(st_epilogue<ST_RESID>(xr, q[j][R], s[j], (RT)w, (RT)sc.c[0]));
                else
                    acc += tnorm2(rp[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < ROWS; j++)
#pragma unroll
            for (int k = 0; k < 2 * R; k++) q[j][k] = q[j][k + 1];
    }
    if (WRITE && i0e > i0s) {
#pragma unroll
        for (int j = 0; j < ROWS; j++)
            if (inr[j]) tile_store<NTS>(o + (i64)(i0e - 1) * plane + pref_off(qr[j]), rp[j]);
    }
    return acc;
}

template <class V, int R, int TZ, int ROWS, int MODE, bool NORM, bool WRITE, int NT>
__global__ __launch_bounds__(NT) void k_star_tile(const V* __restrict__ in, V* __restrict__ out,
                                                   const V* __restrict__ aux, Geom g, Phases ph,
                                                   TileCoefs sc, double w, TileShape sh, i64 ng,
                                                   double* __restrict__ partial)
{
    typedef TileGeom<V, R, TZ, ROWS, NT> TG;
    __shared__ V lds[TG::LDS];
    const int b = blockIdx.x;
    const int t = xcd_tile(b, sh.nblocks);
    const int tz = t % sh.ntz;
    int rest = t / sh.ntz;
    const int ty = rest % sh.nty;
    rest /= sh.nty;
    const int ch = rest % sh.nchunk;
    const int band = rest / sh.nchunk;
    const i64 boff = (i64)band * ng;
    const int i0s = ch * sh.q;
    const V* xa = (aux && MODE != ST_HPSI) ? aux + boff : aux;
    const double acc = star_tile_sweep<V, R, TZ, ROWS, MODE, NORM, WRITE, NT>(
        in + boff, out + boff, xa, g, ph, sc, w, tz * TZ, ty * TG::TY, i0s,
        min(i0s + sh.q, g.n[0]), lds);
    if (NORM) {
        __shared__ double red[NT];
        const double sum = block_sum<NT>(acc, red);
        if (threadIdx.x == 0 && threadIdx.y == 0) partial[b] = sum;
    }
}

// Rows per thread.  One row keeps the register queue small (R=3: 58 VGPRs
// for float64, 118 for complex128, 8 and 4 waves per SIMD); on MI355X the
// occupancy gained outweighs the extra L2 reads of the taller halo: 1 row
// measured 4.1 / 3.6 TB/s on 64^3 float64 / complex128 bands against
// 2.7-3.3 / 3.0-3.2 TB/s for 2-4 rows.
#ifndef GPAW_TILE_ROWS_F32
#define GPAW_TILE_ROWS_F32 1
#endif
constexpr int tile_rows(int, int esize, int nt = TILE_NT)
{
    return (esize == 4 && nt == TILE_NT) ? GPAW_TILE_ROWS_F32 : 1;
}

// host-side shape selection and launch (stencil_tile_*.hip)
int tile_tz(int n2);
// tiles of TZ columns and q planes; false if a tile's halo would wrap twice
bool tile_geometry(const Geom& g, int R, int esize, int nin, int TZ, int q, TileShape& sh,
                   int nt = TILE_NT);
int tile_default_q(const Geom& g, int R, int esize, int nin, int TZ, int nt = TILE_NT);
bool tile_shape(const Geom& g, int R, int esize, int nin, int& TZ, TileShape& sh,
                int nt = TILE_NT);
hipError_t launch_stencil_tile(const Stencil& s, int dtype, int mode, const void* in, void* out,
                               const void* aux, const Geom& g, const Phases& ph, int nin,
                               double w, double* partial, int* npartial, bool write,
                               hipStream_t st, bool* handled);

}  // namespace gpaw
