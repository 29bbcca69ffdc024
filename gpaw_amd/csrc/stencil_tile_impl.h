// Host launcher of k_star_tile for one storage type (instantiated once per
// type in stencil_tile_<type>.hip so the variants compile in parallel).
#pragma once
#include "stencil_tile.h"

namespace gpaw {

template <class V, int R, int TZ, int MODE, bool NORM, bool WRITE>
hipError_t tile_go(const Stencil& s, const void* in, void* out, const void* aux, const Geom& g,
                   const Phases& ph, double w, const TileShape& sh, double* partial,
                   hipStream_t st)
{
    constexpr int ROWS = tile_rows(R, (int)sizeof(V));
    TileCoefs sc;
    for (int c = 0; c < 6 * R + 1; c++) sc.c[c] = s.coefs[c];
    const i64 ng = (i64)g.n[0] * g.n[1] * g.n[2];
    if constexpr (VT<V>::cplx && R <= 3) {   // wider blocks: complex data, range <= 3
        if (sh.nt == TILE_NT_MID)
            return launch_kernel((k_star_tile<V, R, TZ, 1, MODE, NORM, WRITE, TILE_NT_MID>),
                                 dim3(sh.nblocks), dim3(TZ, TILE_NT_MID / TZ), 0, st,
                                 (const V*)in, (V*)out, (const V*)aux, g, ph, sc, w, sh, ng,
                                 partial);
        if (sh.nt == TILE_NT_WIDE)
            return launch_kernel((k_star_tile<V, R, TZ, tile_rows(R, (int)sizeof(V), TILE_NT_WIDE),
                                              MODE, NORM, WRITE, TILE_NT_WIDE>),
                                 dim3(sh.nblocks), dim3(TZ, TILE_NT_WIDE / TZ), 0, st,
                                 (const V*)in, (V*)out, (const V*)aux, g, ph, sc, w, sh, ng,
                                 partial);
    }
    if (sh.nt != TILE_NT) return hipErrorInvalidValue;
    launch_kernel((k_star_tile<V, R, TZ, ROWS, MODE, NORM, WRITE, TILE_NT>), dim3(sh.nblocks),
                  dim3(TZ, TILE_NT / TZ), 0, st, (const V*)in, (V*)out, (const V*)aux, g, ph, sc,
                  w, sh, ng, partial);
    return launch_error();
}

template <class V, int R, int TZ, int MODE>
hipError_t tile_mode(const Stencil& s, const void* in, void* out, const void* aux, const Geom& g,
                     const Phases& ph, double w, const TileShape& sh, double* partial, bool write,
                     hipStream_t st)
{
    if constexpr (!st_mode_ok<V, MODE>()) {
        return hipErrorInvalidValue;
    } else {
        if constexpr (st_norm_ok<V, MODE>()) {
            if (partial && write)
                return tile_go<V, R, TZ, MODE, true, true>(s, in, out, aux, g, ph, w, sh, partial, st);
            if (partial)
                return tile_go<V, R, TZ, MODE, true, false>(s, in, out, aux, g, ph, w, sh, partial, st);
        }
        if (partial) return hipErrorInvalidValue;
        return tile_go<V, R, TZ, MODE, false, true>(s, in, out, aux, g, ph, w, sh, nullptr, st);
    }
}

template <class V, int R, int TZ>
hipError_t tile_modes(const Stencil& s, int mode, const void* in, void* out, const void* aux,
                      const Geom& g, const Phases& ph, double w, const TileShape& sh,
                      double* partial, bool write, hipStream_t st)
{
    // partials: residual norms (RESID; JACOBI: of the input, the previous
    // V-cycle's error) or in . out (APPLY, the mixed solver's CG step);
    // other modes never carry them
    if (partial && mode != ST_RESID && mode != ST_APPLY && mode != ST_JACOBI)
        return hipErrorInvalidValue;
    switch (mode) {
    case ST_APPLY: return tile_mode<V, R, TZ, ST_APPLY>(s, in, out, aux, g, ph, w, sh, partial, true, st);
    case ST_RESID: return tile_mode<V, R, TZ, ST_RESID>(s, in, out, aux, g, ph, w, sh, partial, write, st);
    case ST_JACOBI: return tile_mode<V, R, TZ, ST_JACOBI>(s, in, out, aux, g, ph, w, sh, partial, true, st);
    case ST_PCUPD: return tile_mode<V, R, TZ, ST_PCUPD>(s, in, out, aux, g, ph, w, sh, nullptr, true, st);
    case ST_PCFIN: return tile_mode<V, R, TZ, ST_PCFIN>(s, in, out, aux, g, ph, w, sh, nullptr, true, st);
    case ST_HPSI: return tile_mode<V, R, TZ, ST_HPSI>(s, in, out, aux, g, ph, w, sh, nullptr, true, st);
    }
    return hipErrorInvalidValue;
}

template <class V>
hipError_t tile_dispatch(const Stencil& s, int mode, const void* in, void* out, const void* aux,
                         const Geom& g, const Phases& ph, double w, int TZ, const TileShape& sh,
                         double* partial, bool write, hipStream_t st)
{
#define GPAW_TILE_R(RR)                                                                      \
    case RR:                                                                                 \
        return TZ == 64 ? tile_modes<V, RR, 64>(s, mode, in, out, aux, g, ph, w, sh, partial, \
                                                 write, st)                                  \
                        : tile_modes<V, RR, 32>(s, mode, in, out, aux, g, ph, w, sh, partial, \
                                                 write, st);
    switch (s.star_r) {
        GPAW_TILE_R(1)
        GPAW_TILE_R(2)
        GPAW_TILE_R(3)
        GPAW_TILE_R(4)
    }
#undef GPAW_TILE_R
    return hipErrorInvalidValue;
}

}  // namespace gpaw
