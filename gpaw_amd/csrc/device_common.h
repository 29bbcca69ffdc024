// Shared device-side helpers for the gfx950 finite-difference kernels.
//
// Data layout: every grid function is a C-contiguous [band][i0][i1][i2]
// array in HBM with i2 fastest (the reference layout, c/bmgs/fd.c:16-51),
// real (double / float) or complex (double2 / float2, interleaved re,im).
// There are no padded halo buffers: every kernel reads the interior array
// and produces halo values on the fly ("ghost indexing"), reproducing the
// values the reference writes into its padded buffer in c/bc.c:138-250:
//   * an index below 0 / at or above n on a zero-boundary axis reads 0;
//   * on a periodic axis it wraps once and is multiplied by the Bloch phase
//     of that side: phase[a][0] for the high halo (data sent from the low
//     edge), phase[a][1] for the low halo;
//   * a point outside on several axes gets the phases applied in axis order
//     0, 1, 2, which is the order in which the reference copies its slabs
//     (corners carry products of phases, c/bc.c:108-112).
// Complex*complex products are evaluated as (ac - bd, ad + bc) and
// complex*real as (a*c, b*c), with no FMA contraction (the library is
// built with -ffp-contract=off) so FP64 results are bit-identical to the
// reference C.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpaw {

typedef long long i64;

// per-axis geometry of a grid function
struct Geom {
    int n[3];      // interior extent
    int per[3];    // 1: periodic self-wrap, 0: zero halo
};

// Bloch phases, [axis][side] as in phase_cd of the reference
struct Phases {
    double re[3][2];
    double im[3][2];
};

template <class V> struct VT;
template <> struct VT<double> { typedef double R; enum { cplx = 0 }; };
template <> struct VT<double2> { typedef double R; enum { cplx = 1 }; };
template <> struct VT<float> { typedef float R; enum { cplx = 0 }; };
template <> struct VT<float2> { typedef float R; enum { cplx = 1 }; };

template <class V> __device__ __forceinline__ V vzero();
template <> __device__ __forceinline__ double vzero<double>() { return 0.0; }
template <> __device__ __forceinline__ float vzero<float>() { return 0.0f; }
template <> __device__ __forceinline__ double2 vzero<double2>() { return make_double2(0.0, 0.0); }
template <> __device__ __forceinline__ float2 vzero<float2>() { return make_float2(0.0f, 0.0f); }

// real part, and a value with the given real part (imaginary part 0)
__device__ __forceinline__ double vreal(double v) { return v; }
__device__ __forceinline__ float vreal(float v) { return v; }
__device__ __forceinline__ double vreal(double2 v) { return v.x; }
__device__ __forceinline__ float vreal(float2 v) { return v.x; }
template <class V> __device__ __forceinline__ V vfromreal(typename VT<V>::R r);
template <> __device__ __forceinline__ double vfromreal<double>(double r) { return r; }
template <> __device__ __forceinline__ float vfromreal<float>(float r) { return r; }
template <> __device__ __forceinline__ double2 vfromreal<double2>(double r) { return make_double2(r, 0.0); }
template <> __device__ __forceinline__ float2 vfromreal<float2>(float r) { return make_float2(r, 0.0f); }

__device__ __forceinline__ double vmul(double a, double c) { return a * c; }
__device__ __forceinline__ float vmul(float a, float c) { return a * c; }
__device__ __forceinline__ double2 vmul(double2 a, double c) { return make_double2(a.x * c, a.y * c); }
__device__ __forceinline__ float2 vmul(float2 a, float c) { return make_float2(a.x * c, a.y * c); }

__device__ __forceinline__ double vadd(double a, double b) { return a + b; }
__device__ __forceinline__ float vadd(float a, float b) { return a + b; }
__device__ __forceinline__ double2 vadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 vadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }

__device__ __forceinline__ double vsub(double a, double b) { return a - b; }
__device__ __forceinline__ float vsub(float a, float b) { return a - b; }
__device__ __forceinline__ double2 vsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 vsub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// Division by a real coefficient.  FP64: a true division (the reference
// divides).  FP32 (only the mixed solver's preconditioning V-cycle and the
// FP32 variants, which are compared with FP64 to a tolerance): multiplication
// by the reciprocal, which the compiler hoists out of the loops.
__device__ __forceinline__ double vdiv_r(double a, double c) { return a / c; }
__device__ __forceinline__ float vdiv_r(float a, float c) { return a * (1.0f / c); }
__device__ __forceinline__ double2 vdiv_r(double2 a, double c) { return make_double2(a.x / c, a.y / c); }
__device__ __forceinline__ float2 vdiv_r(float2 a, float c)
{
    const float r = 1.0f / c;
    return make_float2(a.x * r, a.y * r);
}

// x + v * c: FP64 rounds the product and the sum separately (the reference's
// x += a * coef, never contracted); FP32 uses one fused multiply-add (half
// the arithmetic of the FP32 stencils, one rounding instead of two).
__device__ __forceinline__ double vmac(double x, double v, double c) { return x + v * c; }
__device__ __forceinline__ double2 vmac(double2 x, double2 v, double c)
{
    return make_double2(x.x + v.x * c, x.y + v.y * c);
}
__device__ __forceinline__ float vmac(float x, float v, float c) { return __builtin_fmaf(v, c, x); }
__device__ __forceinline__ float2 vmac(float2 x, float2 v, float c)
{
    return make_float2(__builtin_fmaf(v.x, c, x.x), __builtin_fmaf(v.y, c, x.y));
}

// multiply by a phase factor (complex only; identity for real data)
__device__ __forceinline__ double pmul(double v, double, double) { return v; }
__device__ __forceinline__ float pmul(float v, double, double) { return v; }
__device__ __forceinline__ double2 pmul(double2 v, double pr, double pi)
{
    return make_double2(pr * v.x - pi * v.y, pr * v.y + pi * v.x);
}
__device__ __forceinline__ float2 pmul(float2 v, double pr, double pi)
{
    const float r = (float)pr, i = (float)pi;
    return make_float2(r * v.x - i * v.y, r * v.y + i * v.x);
}

// Multiply by the phase of one side (f = 1: high halo, 2: low halo) of one
// axis.  SEL chooses the side by a select instead of an index: phases held
// in a copy of a struct read from global memory would otherwise be indexed
// at run time and placed in private (scratch) memory; phases passed as
// kernel arguments are indexed in place, which needs fewer registers.
template <bool SEL = false, class V>
__device__ __forceinline__ V pside(V v, const Phases& ph, int axis, int f)
{
    if constexpr (SEL) {
        const double pr = f == 1 ? ph.re[axis][0] : ph.re[axis][1];
        const double pi = f == 1 ? ph.im[axis][0] : ph.im[axis][1];
        return pmul(v, pr, pi);
    } else {
        return pmul(v, ph.re[axis][f - 1], ph.im[axis][f - 1]);
    }
}

// Wrap one index.  Returns 0 = inside, 1 = high halo (phase side 0),
// 2 = low halo (phase side 1), 3 = outside a zero-boundary axis.
__device__ __forceinline__ int wrap_index(int& j, int n, int per)
{
    if (j < 0) {
        if (!per) return 3;
        j += n;
        return 2;
    }
    if (j >= n) {
        if (!per) return 3;
        j -= n;
        return 1;
    }
    return 0;
}

// Value of the padded field at (j0, j1, j2) of one band.
template <class V, bool SEL = false>
__device__ __forceinline__ V ghost_fetch(const V* __restrict__ a, int j0, int j1, int j2,
                                         const Geom& g, const Phases& ph)
{
    const int f0 = wrap_index(j0, g.n[0], g.per[0]);
    const int f1 = wrap_index(j1, g.n[1], g.per[1]);
    const int f2 = wrap_index(j2, g.n[2], g.per[2]);
    if (f0 == 3 || f1 == 3 || f2 == 3) return vzero<V>();
    V v = a[((i64)j0 * g.n[1] + j1) * g.n[2] + j2];
    if (VT<V>::cplx) {
        if (f0) v = pside<SEL>(v, ph, 0, f0);
        if (f1) v = pside<SEL>(v, ph, 1, f1);
        if (f2) v = pside<SEL>(v, ph, 2, f2);
    }
    return v;
}

// In-plane part of a ghost reference, packed in one int: the wrapped offset
// within a plane in bits 4..30 (planes of fewer than 2^27 points) and the
// wrap flags of axes 1 and 2 (wrap_index codes) in bits 0..1 and 2..3.  A
// reference to a zero-boundary ghost has offset 0 (any valid address: the
// value is discarded).
typedef int PlaneRef;
constexpr long PLANE_REF_MAX = 1L << 27;   // points per plane

__device__ __forceinline__ PlaneRef plane_ref(int j1, int j2, const Geom& g)
{
    const int f1 = wrap_index(j1, g.n[1], g.per[1]);
    const int f2 = wrap_index(j2, g.n[2], g.per[2]);
    const int off = (f1 == 3 || f2 == 3) ? 0 : j1 * g.n[2] + j2;
    return (off << 4) | f1 | (f2 << 2);
}
__device__ __forceinline__ int pref_off(PlaneRef r) { return r >> 4; }
constexpr PlaneRef PREF_ZERO = 3;   // offset 0, axis-1 flag "zero"

// Value of a ghost reference from its raw load: zero outside a zero-boundary
// axis, else the phases of the wrapped sides applied in axis order 0, 1, 2
// (the order of ghost_fetch).  Phase sides are selected, not indexed, so the
// phases stay in scalar registers.
template <class V>
__device__ __forceinline__ V ghost_fix(V v, int f0, PlaneRef r, const Phases& ph)
{
    const int f1 = r & 3, f2 = (r >> 2) & 3;
    if (f0 == 3 || f1 == 3 || f2 == 3) return vzero<V>();
    if (VT<V>::cplx) {
        // real branches, not selects (the empty asm keeps the compiler from
        // computing both sides): phases apply on few planes and lanes
        if (f0) {
            asm volatile("");
            v = pside<true>(v, ph, 0, f0);
        }
        if (r & 15) {
            asm volatile("");
            if (f1) v = pside<true>(v, ph, 1, f1);
            if (f2) v = pside<true>(v, ph, 2, f2);
        }
    }
    return v;
}

// Make a loaded value available at this point (an empty asm that reads it:
// the compiler places the wait for the load before it; no instruction).
__device__ __forceinline__ void vtouch(double v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void vtouch(float v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void vtouch(double2 v) { asm volatile("" ::"v"(v.x), "v"(v.y)); }
__device__ __forceinline__ void vtouch(float2 v) { asm volatile("" ::"v"(v.x), "v"(v.y)); }

// Tile index of block b of nb: block b runs on XCD b % 8, so each XCD gets
// a contiguous range of tiles and neighbouring tiles, which share halo
// rows, meet in one L2.  A bijection on [0, nb).
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    return (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3);
}

// Deterministic block reduction of one double per thread (fixed tree order).
// Every thread of the block must call it.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* lds)
{
    const int t = threadIdx.x + threadIdx.y * blockDim.x + threadIdx.z * blockDim.x * blockDim.y;
    lds[t] = v;
    __syncthreads();
#pragma unroll
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (t < s) lds[t] = lds[t] + lds[t + s];
        __syncthreads();
    }
    const double r = lds[0];
    __syncthreads();
    return r;
}

}  // namespace gpaw
