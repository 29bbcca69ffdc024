// Host-side launch interface of the gfx950 kernels (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include "device_common.h"
#include "launch.h"

namespace gpaw {

enum DType { DT_F64 = 0, DT_C128 = 1, DT_F32 = 2, DT_C64 = 3 };

inline int dtype_size(int dt) { return dt == DT_F64 ? 8 : dt == DT_C128 ? 16 : dt == DT_F32 ? 4 : 8; }
inline bool dtype_complex(int dt) { return dt == DT_C128 || dt == DT_C64; }

enum StencilMode {
    ST_APPLY = 0,    // out = sum_c coef_c * in[g + off_c]
    ST_RESID = 1,    // out = apply(in) - aux
    ST_JACOBI = 2,   // out = (1-w) in + w (aux - sum_{c>=1} ...) / coef_0
    ST_PCUPD = 3,    // out = in - w (apply(in) - aux)     (FP64 types only)
    ST_PCFIN = 4,    // out = in - w (apply(in) + aux)     (FP64 types only)
    ST_HPSI = 5,     // out = apply(in) + in * v, v a real grid function shared
                     // by all bands, passed as aux      (FP64 types only)
};

// Per-point result of a stencil pass from the stencil sum x (over c >= 1 for
// ST_JACOBI, over all c otherwise), the centre input value f, the aux value
// s, the weight w and the centre coefficient c0.  Each operation is rounded
// on its own, in the order the reference evaluates it:
//   ST_JACOBI  c/bmgs/relax.c:60;
//   ST_PCUPD   q = L d; q -= r; d -= w * q         (gpaw/preconditioner.py:56-66);
//   ST_PCFIN   q = L d; q -= res; d += -w * q; d *= -1 evaluated on e = -d
//              (gpaw/preconditioner.py:67-70: negation commutes exactly with
//              the linear stencil and with rounding, so the value equals
//              e - w * (L e + res));
//   ST_HPSI    Htpsi = kin psi; Htpsi += psi * vt (gpaw/wavefunctions/fd.py:51-52,
//              gpaw/hamiltonian.py:393-398); s carries vt in its real part.
template <int MODE, class V, class RT>
__device__ __forceinline__ V st_epilogue(V x, V f, V s, RT w, RT c0)
{
    if (MODE == ST_APPLY) return x;
    if (MODE == ST_RESID) return vsub(x, s);
    if (MODE == ST_JACOBI) return vadd(vmul(f, (RT)1 - w), vdiv_r(vmul(vsub(s, x), w), c0));
    if (MODE == ST_PCUPD) return vsub(f, vmul(vsub(x, s), w));
    if (MODE == ST_HPSI) return vadd(x, vmul(f, vreal(s)));
    return vsub(f, vmul(vadd(x, s), w));
}

// The aux value of point `idx` of a band whose first point is `boff`: a
// band's own aux array, or for ST_HPSI the real potential shared by all
// bands (its value in the real part).
template <int MODE, class V>
__device__ __forceinline__ V aux_value(const V* __restrict__ aux, i64 boff, i64 idx)
{
    if (MODE == ST_HPSI) return vfromreal<V>(((const typename VT<V>::R*)aux)[idx]);
    return aux[boff + idx];
}

// Which (type, mode, norm) combinations are compiled.
template <class V, int MODE>
constexpr bool st_mode_ok()
{
    if (MODE == ST_PCUPD || MODE == ST_PCFIN || MODE == ST_HPSI)
        return sizeof(typename VT<V>::R) == 8;
    if (MODE == ST_JACOBI) return !VT<V>::cplx;
    return true;
}
template <class V, int MODE>
constexpr bool st_norm_ok()
{
    return !VT<V>::cplx && (MODE == ST_APPLY || MODE == ST_RESID || MODE == ST_JACOBI);
}

#define GPAW_MAXC 512
#define GPAW_MAXR 6

// A stencil: coefficients in the caller's order (the summation order) and
// integer offsets.  `star` marks the canonical order of an orthogonal-cell
// Laplacian, [centre, +1..+R along axis 2, -1..-R along 2, then axis 1,
// then axis 0], for which the register-queue kernel is used.
struct Stencil {
    int nc = 0;
    int range = 0;
    int star_r = 0;                 // R of the canonical star order, 0 if not
    double coefs[GPAW_MAXC];
    int off[GPAW_MAXC][3];
    double* d_coefs = nullptr;      // device copies for the generic kernel
    int* d_off = nullptr;
};

// Stencil launch-shape autotuner (stencil.hip): switch on (1) / off (0), or
// query (-1); returns the previous state.  autotune_report writes a text
// table of the choices made.
int autotune_set(int on);
// Kernel family of star-stencil launches without partial sums: -1 default,
// 0 generic, 1 register queue, 2 tile, 3 box; returns the previous value.
int stencil_kernel_set(int kind);
int tile_shape_set(int nt, int tz, int q);   // gpaw_tile_shape
long autotune_report(char* buf, long len);

// Work estimate used to size partial-sum buffers.
int stencil_partials(const Stencil& s, const Geom& g, int nin);
// Whether an ST_APPLY pass of this stencil and grid can also produce the
// partial sums of in . out, and an ST_JACOBI sweep those of |L in - aux|^2
// (real data on the tile kernel; the same per-block layout as a residual
// norm pass); launch_stencil refuses these partials otherwise.
bool stencil_dot_ok(const Stencil& s, int dtype, const Geom& g, int nin);

// S = JF_SWEEPS weighted-Jacobi sweeps of a 7-point star stencil in one
// launch (temporal blocking, jacobi_fused.hip): real data, in != out;
// bit-identical to S single sweeps ping-ponging between two arrays, with
// the result in `out`.
struct StarCoefs7 {
    double c[7];
};
#define JF_SWEEPS 4
// With `corr` (the coarser level's result) and `ctr` (the K = 2
// interpolator from it onto this grid), the loaded field is first corrected,
// in -= I corr, as a separate interpolate-and-subtract launch would.
// `zero_in`: the caller guarantees in == 0 everywhere (a level's first
// sweeps after the restriction zeroed it), so it is not read.
struct Transfer;
// jacobi_fused_sweeps: sweeps per launch on this grid (JF_SWEEPS, or 2 on
// grids large enough for 12^3 output boxes to fill the chip); `sweeps` of
// launch_jacobi_fused is 2 or 4.
bool jacobi_fused_ok(const Stencil& s, const Geom& g);
int jacobi_fused_sweeps(const Geom& g);
// LDS box edge of a fused launch of `sweeps` sweeps on g (16, or 12 on grids
// with too few 16^3 boxes; GPAW_AMD_JF_BOX=16 or 12 forces one)
int jacobi_fused_edge(const Geom& g, int sweeps);
// two Jacobi sweeps from phi = 0 (weights w0 then w1) in one box launch, the
// bits of k_jacobi_from_zero followed by one sweep; real FP64 / FP32, star
// range 1..3 (hipErrorInvalidValue otherwise: use the two launches)
bool jacobi_two_from_zero_ok(const Stencil& s, int dtype, const Geom& g);
hipError_t launch_jacobi_two_from_zero(const Stencil& s, int dtype, const void* rho, void* out,
                                       const Geom& g, double w0, double w1, hipStream_t st);
hipError_t launch_jacobi_fused(const Stencil& s, int dtype, const void* in, void* out,
                               const void* rho, const Geom& g, double w, hipStream_t st,
                               const void* corr = nullptr, const Transfer* ctr = nullptr,
                               bool zero_in = false, int sweeps = JF_SWEEPS,
                               const double* ws = nullptr);   // per-sweep weights (else w)

// Residual of a 7-point star stencil and its K = 2 restriction in one
// launch: rho_c = restrict(L phi - rho), and phi_c = 0 when given; the
// same values as an ST_RESID pass followed by a TR_STORE restriction.
bool resid_restrict_ok(const Stencil& s, const Geom& g, const Transfer& tr);
hipError_t launch_resid_restrict(const Stencil& s, int dtype, const void* phi, const void* rho,
                                 void* rho_c, void* phi_c, const Geom& g, const Transfer& tr,
                                 hipStream_t st);

// One weighted-Jacobi sweep of a star stencil (R = 1 .. 3, real data) whose
// input is first corrected by phi -= I e (K = 2 interpolation `itp` of the
// coarser level's e onto g): the same values as the interpolate-and-subtract
// launch followed by the sweep.  `in` itself is not modified.
bool jacobi_corr_ok(const Stencil& s, const Geom& g, const Transfer& itp);
hipError_t launch_jacobi_corr(const Stencil& s, int dtype, const void* in, void* out,
                              const void* rho, const void* corr, const Geom& g, double w,
                              const Transfer& itp, hipStream_t st);

// Launch one stencil pass.  `partial` (optional) receives one FP64 partial
// sum of |result|^2 per thread block; `write` = false skips the store.
hipError_t launch_stencil(const Stencil& s, int dtype, int mode, const void* in, void* out,
                          const void* aux, const Geom& g, const Phases& ph, int nin,
                          double w, double* partial, int* npartial, bool write,
                          hipStream_t st);

// Restriction / interpolation.
struct Transfer {
    int K = 2;
    int interpolate = 0;
    Geom gin;            // input geometry (n, periodic flags)
    int nout[3];
    int lo[3];           // low padding per axis (maps output to input index)
    int skip[3][2];
};

enum TransferMode {
    TR_STORE = 0,        // out = T(in)
    TR_SUBTRACT = 1,     // out -= T(in)   (interpolate + correct)
};

hipError_t launch_transfer(const Transfer& t, int dtype, int mode, const void* in, void* out,
                           void* zero_out, const Phases& ph, int nin, hipStream_t st);

// Reductions and element-wise helpers (FP64 real unless stated).
hipError_t launch_reduce_partials(const double* partial, int n, double scale, double* result,
                                  hipStream_t st);
// result[i] = scale * sum_j partial[i * ncols + j], each row exactly as
// launch_reduce_partials sums it
hipError_t launch_reduce_rows(const double* partial, int nrows, int ncols, double scale,
                              double* result, hipStream_t st);
hipError_t launch_sum(const double* a, long long n, double* partial, int* npartial,
                      hipStream_t st);
// result[j] = scale * sum_g a[g] * b[j n + g] for j < nb (nb <= 65535);
// partial holds nb * dots_chunks(n) doubles.
int dots_chunks(long long n);
hipError_t launch_dots(const double* a, const double* b, int nb, long long n, double scale,
                       double* partial, double* result, hipStream_t st);
hipError_t launch_axpby(double* y, const double* x, double a, double b, long long n,
                        hipStream_t st);   // y = a*y + b*x
hipError_t launch_sub_scalar(double* y, const double* x, double b, long long n,
                             hipStream_t st);  // y = x - b
hipError_t launch_fill(void* y, int dtype, long long n, hipStream_t st);
hipError_t launch_copy(void* y, const void* x, int dtype, long long n, hipStream_t st);
hipError_t launch_cast(void* y, int dty, const void* x, int dtx, long long n, hipStream_t st);
// nbytes a multiple of 16, both pointers 16-byte aligned
hipError_t launch_copy_bytes(void* y, const void* x, long long nbytes, hipStream_t st);
hipError_t launch_add_potential(void* out, const void* in, const double* v, long long nin,
                                long long ng, int dtype, hipStream_t st);
hipError_t launch_scale(double* y, double a, long long n, hipStream_t st);

}  // namespace gpaw
