// Batched star-stencil apply: many grid functions of different shapes,
// coefficients and Bloch phases (bands of several k-points, independent
// grids of uneven size) in one persistent launch per stencil range.
//
// Each group (one operator applied to nin bands of one grid) is cut into
// the tiles of the single-grid tile kernel; the tiles of all groups form
// one list.  A grid of co-resident blocks walks the list in rounds: in
// round r, block b takes tile r * nblocks + xcd_tile(b), so every XCD works
// on a contiguous range of tiles and the tiles of one band meet in one L2.
// The group of a tile is found by a binary search over the groups' first
// tile indices (uniform per block, scalar loads).  Per-point arithmetic is
// the tile kernel's, so results are bit-identical to single applies.
#pragma once
#include "stencil_tile.h"

namespace gpaw {

struct BatchGroup {
    const void* in;
    void* out;
    Geom g;
    Phases ph;
    TileCoefs sc;
    TileShape sh;
    i64 ng;        // points per band
    int first;     // index of the group's first tile in the list
    int pad;
};

// NT threads per block: 256, or 1024 for complex data of range <= 3 (the
// single-grid kernel's wide tiles: taller tiles, fewer halo rows)
template <class V, int R, int TZ, int NT>
__global__ __launch_bounds__(NT) void k_star_tile_batch(const BatchGroup* __restrict__ groups,
                                                        int ngroups, int ntiles)
{
    constexpr int ROWS = tile_rows(R, (int)sizeof(V), NT);
    typedef TileGeom<V, R, TZ, ROWS, NT> TG;
    __shared__ V lds[TG::LDS];
    const int nb = gridDim.x;
    const int slot = xcd_tile(blockIdx.x, nb);
    for (int base = 0; base + slot < ntiles; base += nb) {
        const int t = base + slot;
        int lo = 0, hi = ngroups - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (groups[mid].first <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        const BatchGroup* gp = groups + lo;
        const TileShape sh = gp->sh;
        int r = t - gp->first;
        const int tz = r % sh.ntz;
        r /= sh.ntz;
        const int ty = r % sh.nty;
        r /= sh.nty;
        const int ch = r % sh.nchunk;
        const int band = r / sh.nchunk;
        const i64 boff = (i64)band * gp->ng;
        const int i0s = ch * sh.q;
        star_tile_sweep<V, R, TZ, ROWS, ST_APPLY, false, true, NT>(
            (const V*)gp->in + boff, (V*)gp->out + boff, nullptr, gp->g, gp->ph, gp->sc, 0.0,
            tz * TZ, ty * TG::TY, i0s, min(i0s + sh.q, gp->g.n[0]), lds);
        __syncthreads();   // the next tile rewrites the LDS plane buffers
    }
}

// Resident-block count of the (V, R, TZ, NT) batch kernel on this device.
template <class V, int R, int TZ, int NT>
int batch_resident()
{
    static int n = 0;
    if (!n) {
        int per_cu = 0, dev = 0, cus = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, (const void*)k_star_tile_batch<V, R, TZ, NT>, NT, 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        n = per_cu * (cus > 0 ? cus : 256);
    }
    return n;
}

template <class V, int R, int TZ, int NT>
hipError_t batch_go_nt(const BatchGroup* d_groups, int ngroups, int ntiles, hipStream_t st)
{
    int nb = batch_resident<V, R, TZ, NT>();
    if (nb >= ntiles)
        nb = ntiles;
    else if (nb >= 8)
        nb &= ~7;   // whole rounds of 8 XCDs
    return launch_kernel(k_star_tile_batch<V, R, TZ, NT>, dim3(nb), dim3(TZ, NT / TZ), 0, st,
                         d_groups, ngroups, ntiles);
}

// block size a batch of range R uses for storage type V (groups are cut
// into tiles of that size)
template <class V>
constexpr int batch_nt(int R)
{
    return (VT<V>::cplx && R <= 3) ? TILE_NT_WIDE : TILE_NT;
}

template <class V, int R, int TZ>
hipError_t batch_go(const BatchGroup* d_groups, int ngroups, int ntiles, int nt, hipStream_t st)
{
    if constexpr (VT<V>::cplx && R <= 3) {
        if (nt == TILE_NT_WIDE)
            return batch_go_nt<V, R, TZ, TILE_NT_WIDE>(d_groups, ngroups, ntiles, st);
    }
    if (nt != TILE_NT) return hipErrorInvalidValue;
    return batch_go_nt<V, R, TZ, TILE_NT>(d_groups, ngroups, ntiles, st);
}

template <class V>
hipError_t batch_dispatch(const BatchGroup* d_groups, int ngroups, int ntiles, int R, int TZ,
                          int nt, hipStream_t st)
{
#define GPAW_BATCH_R(RR)                                                                  \
    case RR:                                                                              \
        return TZ == 64 ? batch_go<V, RR, 64>(d_groups, ngroups, ntiles, nt, st)         \
                        : batch_go<V, RR, 32>(d_groups, ngroups, ntiles, nt, st);
    switch (R) {
        GPAW_BATCH_R(1)
        GPAW_BATCH_R(2)
        GPAW_BATCH_R(3)
        GPAW_BATCH_R(4)
    }
#undef GPAW_BATCH_R
    return hipErrorInvalidValue;
}

hipError_t batch_launch(int dtype, const BatchGroup* d_groups, int ngroups, int ntiles, int R,
                        int TZ, int nt, hipStream_t st);

}  // namespace gpaw
