// Host launcher of k_star_box for one storage type (instantiated once per
// type in stencil_box_<type>.hip so the variants compile in parallel).
#pragma once
#include "stencil_box.h"

namespace gpaw {

template <class V, int R, int Q, int MODE, bool NORM, bool WRITE>
hipError_t box_go(const Stencil& s, const void* in, void* out, const void* aux, const Geom& g,
                  const Phases& ph, double w, const TileShape& sh, double* partial,
                  hipStream_t st)
{
    TileCoefs sc;
    for (int c = 0; c < 6 * R + 1; c++) sc.c[c] = s.coefs[c];
    const i64 ng = (i64)g.n[0] * g.n[1] * g.n[2];
    launch_kernel((k_star_box<V, R, Q, MODE, NORM, WRITE>), dim3(sh.nblocks),
                  dim3(BOX_TZ, BOX_NTY), 0, st, (const V*)in, (V*)out, (const V*)aux, g, ph, sc, w,
                  sh, ng, partial, 0.0);
    return launch_error();
}

// two Jacobi sweeps from phi = 0 (weights w0, w1) in one launch (FZ)
template <class V, int R, int Q>
hipError_t box_fz_go(const Stencil& s, const void* rho, void* out, const Geom& g, double w0,
                     double w1, const TileShape& sh, hipStream_t st)
{
    TileCoefs sc;
    for (int c = 0; c < 6 * R + 1; c++) sc.c[c] = s.coefs[c];
    const i64 ng = (i64)g.n[0] * g.n[1] * g.n[2];
    Phases ph;   // real data: no phases
    for (int a = 0; a < 3; a++)
        for (int e = 0; e < 2; e++) {
            ph.re[a][e] = 1.0;
            ph.im[a][e] = 0.0;
        }
    launch_kernel((k_star_box<V, R, Q, ST_JACOBI, false, true, true>), dim3(sh.nblocks),
                  dim3(BOX_TZ, BOX_NTY), 0, st, (const V*)rho, (V*)out, (const V*)rho, g, ph, sc,
                  w1, sh, ng, (double*)nullptr, w0);
    return launch_error();
}

template <class V>
hipError_t box_fz(const Stencil& s, const void* rho, void* out, const Geom& g, double w0,
                  double w1, const TileShape& sh, hipStream_t st)
{
#define GPAW_FZ_Q(R)                                                              \
    switch (sh.q) {                                                               \
    case 2: return box_fz_go<V, R, 2>(s, rho, out, g, w0, w1, sh, st);            \
    case 4: return box_fz_go<V, R, 4>(s, rho, out, g, w0, w1, sh, st);            \
    case 8: return box_fz_go<V, R, 8>(s, rho, out, g, w0, w1, sh, st);            \
    }                                                                             \
    return hipErrorInvalidValue;
    switch (s.star_r) {
    case 1: GPAW_FZ_Q(1)
    case 2: GPAW_FZ_Q(2)
    case 3: GPAW_FZ_Q(3)
    }
#undef GPAW_FZ_Q
    return hipErrorInvalidValue;
}

template <class V, int R, int Q, int MODE>
hipError_t box_mode(const Stencil& s, const void* in, void* out, const void* aux, const Geom& g,
                    const Phases& ph, double w, const TileShape& sh, double* partial, bool write,
                    hipStream_t st)
{
    if constexpr (!st_mode_ok<V, MODE>()) {
        return hipErrorInvalidValue;
    } else {
        if constexpr (st_norm_ok<V, MODE>()) {
            if (partial && write)
                return box_go<V, R, Q, MODE, true, true>(s, in, out, aux, g, ph, w, sh, partial, st);
            if (partial && MODE == ST_RESID)
                return box_go<V, R, Q, MODE, true, false>(s, in, out, aux, g, ph, w, sh, partial, st);
        }
        if (partial) return hipErrorInvalidValue;
        return box_go<V, R, Q, MODE, false, true>(s, in, out, aux, g, ph, w, sh, nullptr, st);
    }
}

template <class V, int R, int Q>
hipError_t box_modes(const Stencil& s, int mode, const void* in, void* out, const void* aux,
                     const Geom& g, const Phases& ph, double w, const TileShape& sh,
                     double* partial, bool write, hipStream_t st)
{
    if (partial && mode != ST_RESID && mode != ST_APPLY && mode != ST_JACOBI)
        return hipErrorInvalidValue;
    switch (mode) {
    case ST_APPLY: return box_mode<V, R, Q, ST_APPLY>(s, in, out, aux, g, ph, w, sh, partial, true, st);
    case ST_RESID: return box_mode<V, R, Q, ST_RESID>(s, in, out, aux, g, ph, w, sh, partial, write, st);
    case ST_JACOBI: return box_mode<V, R, Q, ST_JACOBI>(s, in, out, aux, g, ph, w, sh, partial, true, st);
    case ST_PCUPD: return box_mode<V, R, Q, ST_PCUPD>(s, in, out, aux, g, ph, w, sh, nullptr, true, st);
    case ST_PCFIN: return box_mode<V, R, Q, ST_PCFIN>(s, in, out, aux, g, ph, w, sh, nullptr, true, st);
    case ST_HPSI: return box_mode<V, R, Q, ST_HPSI>(s, in, out, aux, g, ph, w, sh, nullptr, true, st);
    }
    return hipErrorInvalidValue;
}

template <class V, int R>
hipError_t box_q(const Stencil& s, int mode, const void* in, void* out, const void* aux,
                 const Geom& g, const Phases& ph, double w, const TileShape& sh,
                 double* partial, bool write, hipStream_t st)
{
    switch (sh.q) {
    case 2: return box_modes<V, R, 2>(s, mode, in, out, aux, g, ph, w, sh, partial, write, st);
    case 4: return box_modes<V, R, 4>(s, mode, in, out, aux, g, ph, w, sh, partial, write, st);
    case 8: return box_modes<V, R, 8>(s, mode, in, out, aux, g, ph, w, sh, partial, write, st);
    }
    return hipErrorInvalidValue;
}

template <class V>
hipError_t box_dispatch(const Stencil& s, int mode, const void* in, void* out, const void* aux,
                        const Geom& g, const Phases& ph, double w, const TileShape& sh,
                        double* partial, bool write, hipStream_t st)
{
    switch (s.star_r) {
    case 1: return box_q<V, 1>(s, mode, in, out, aux, g, ph, w, sh, partial, write, st);
    case 2: return box_q<V, 2>(s, mode, in, out, aux, g, ph, w, sh, partial, write, st);
    case 3: return box_q<V, 3>(s, mode, in, out, aux, g, ph, w, sh, partial, write, st);
    case 4: return box_q<V, 4>(s, mode, in, out, aux, g, ph, w, sh, partial, write, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace gpaw
