// Reductions and element-wise kernels used around the stencils
// (gfx950).  FP64 sums are deterministic: one partial per block in a fixed
// tree order, then one block sums the partials in index order.
#include "kernels.h"

namespace gpaw {

GraphRecorder*& graph_recorder()
{
    static thread_local GraphRecorder* rec = nullptr;
    return rec;
}

__global__ __launch_bounds__(256) void k_reduce_partials(const double* __restrict__ p, int n,
                                                         double scale, double* __restrict__ out)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += p[i];
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) out[0] = s * scale;
}

__global__ __launch_bounds__(256) void k_sum(const double* __restrict__ a, i64 n,
                                             double* __restrict__ partial)
{
    __shared__ double red[256];
    double s = 0.0;
    const i64 stride = (i64)gridDim.x * 256;
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) s += a[i];
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Row j of a block of dot products: partial sums of a[g] * b[j n + g] over
// chunk blockIdx.x of the grid (the product rounded on its own, then added,
// as numpy's rhot * phit followed by a sum).
__global__ __launch_bounds__(256) void k_dots(const double* __restrict__ a,
                                              const double* __restrict__ b, i64 n, int nchunk,
                                              double* __restrict__ partial)
{
    __shared__ double red[256];
    const double* __restrict__ bj = b + (i64)blockIdx.y * n;
    double s = 0.0;
    const i64 stride = (i64)nchunk * 256;
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) s += a[i] * bj[i];
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) partial[(i64)blockIdx.y * nchunk + blockIdx.x] = s;
}

// out[j] = scale * (sum of row j's partials, in index order); row by row
// the same order and tree as k_reduce_partials
__global__ __launch_bounds__(256) void k_reduce_rows(const double* __restrict__ p, int nchunk,
                                                     double scale, double* __restrict__ out)
{
    __shared__ double red[256];
    const double* __restrict__ pj = p + (i64)blockIdx.x * nchunk;
    double s = 0.0;
    for (int i = threadIdx.x; i < nchunk; i += 256) s += pj[i];
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) out[blockIdx.x] = s * scale;
}

__global__ __launch_bounds__(256) void k_axpby(double* __restrict__ y, const double* __restrict__ x,
                                               double a, double b, i64 n)
{
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = a * y[i] + b * x[i];
}

// y = x - b  (numpy's rho - background, phi -= phi_ave)
__global__ __launch_bounds__(256) void k_sub_scalar(double* __restrict__ y,
                                                    const double* __restrict__ x, double b, i64 n)
{
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = x[i] - b;
}

template <class T>
__global__ __launch_bounds__(256) void k_fill(T* __restrict__ y, i64 n)
{
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = T(0);
}

template <class T>
__global__ __launch_bounds__(256) void k_copy(T* __restrict__ y, const T* __restrict__ x, i64 n)
{
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = x[i];
}

// out[b][g] += in[b][g] * v[g]: the local potential applied to bands
// (numpy's Htpsit_G += psit_G * vt_G, gpaw/hamiltonian.py:393-398); v is
// real and shared by all bands, complex bands scale both parts
template <class V>
__global__ __launch_bounds__(256) void k_add_pot(V* __restrict__ out, const V* __restrict__ in,
                                                 const double* __restrict__ v, i64 ng, i64 n)
{
    typedef typename VT<V>::R RT;
    const i64 stride = (i64)gridDim.x * 256;
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
        out[i] = vadd(out[i], vmul(in[i], (RT)v[i % ng]));
}

// y *= a (numpy's in-place scaling, e.g. nt_sg *= x, gpaw/density.py:234)
__global__ __launch_bounds__(256) void k_scale(double* __restrict__ y, double a, i64 n)
{
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = y[i] * a;
}

// 16-byte copy, four loads in flight per thread, grid-stride over blocks of
// 4 x 256 vectors (the attainable HBM read + write rate)
__global__ __launch_bounds__(256) void k_copy16_flat(uint4* __restrict__ y,
                                                     const uint4* __restrict__ x, i64 n)
{
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = x[i];
}

__global__ __launch_bounds__(256) void k_copy16(uint4* __restrict__ y, const uint4* __restrict__ x,
                                                i64 n)
{
    const i64 stride = (i64)gridDim.x * 1024;
    for (i64 base = (i64)blockIdx.x * 1024 + threadIdx.x; base < n; base += stride) {
        uint4 v[4];
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (base + 256 * k < n) v[k] = x[base + 256 * k];
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (base + 256 * k < n) y[base + 256 * k] = v[k];
    }
}

template <class TO, class TI>
__global__ __launch_bounds__(256) void k_cast(TO* __restrict__ y, const TI* __restrict__ x, i64 n)
{
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = (TO)x[i];
}

static dim3 grid1(i64 n) { return dim3((unsigned)((n + 255) / 256)); }

hipError_t launch_reduce_partials(const double* partial, int n, double scale, double* result,
                                  hipStream_t st)
{
    launch_kernel(k_reduce_partials, dim3(1), dim3(256), 0, st, partial, n, scale, result);
    return launch_error();
}

hipError_t launch_sum(const double* a, long long n, double* partial, int* npartial, hipStream_t st)
{
    int nb = (int)((n + 255) / 256);
    if (nb > 1024) nb = 1024;
    if (nb < 1) nb = 1;
    *npartial = nb;
    launch_kernel(k_sum, dim3(nb), dim3(256), 0, st, a, (i64)n, partial);
    return launch_error();
}

int dots_chunks(long long n)
{
    long long c = (n + 1023) / 1024;
    return (int)(c < 1 ? 1 : c > 256 ? 256 : c);
}

hipError_t launch_reduce_rows(const double* partial, int nrows, int ncols, double scale,
                              double* result, hipStream_t st)
{
    if (nrows <= 0) return hipSuccess;
    launch_kernel(k_reduce_rows, dim3(nrows), dim3(256), 0, st, partial, ncols, scale, result);
    return launch_error();
}

hipError_t launch_dots(const double* a, const double* b, int nb, long long n, double scale,
                       double* partial, double* result, hipStream_t st)
{
    if (nb <= 0) return hipSuccess;
    const int nchunk = dots_chunks(n);
    launch_kernel(k_dots, dim3(nchunk, nb), dim3(256), 0, st, a, b, (i64)n, nchunk, partial);
    launch_kernel(k_reduce_rows, dim3(nb), dim3(256), 0, st, (const double*)partial, nchunk,
                  scale, result);
    return launch_error();
}

hipError_t launch_axpby(double* y, const double* x, double a, double b, long long n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    launch_kernel(k_axpby, grid1(n), dim3(256), 0, st, y, x, a, b, (i64)n);
    return launch_error();
}

hipError_t launch_sub_scalar(double* y, const double* x, double b, long long n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    launch_kernel(k_sub_scalar, grid1(n), dim3(256), 0, st, y, x, b, (i64)n);
    return launch_error();
}

hipError_t launch_fill(void* y, int dtype, long long n, hipStream_t st)
{
    // complex arrays are filled as twice as many reals
    i64 m = (i64)n * (dtype_complex(dtype) ? 2 : 1);
    if (m <= 0) return hipSuccess;
    if (dtype == DT_F64 || dtype == DT_C128)
        launch_kernel(k_fill<double>, grid1(m), dim3(256), 0, st, (double*)y, m);
    else
        launch_kernel(k_fill<float>, grid1(m), dim3(256), 0, st, (float*)y, m);
    return launch_error();
}

hipError_t launch_copy(void* y, const void* x, int dtype, long long n, hipStream_t st)
{
    i64 m = (i64)n * (dtype_complex(dtype) ? 2 : 1);
    if (m <= 0) return hipSuccess;
    if (dtype == DT_F64 || dtype == DT_C128)
        launch_kernel(k_copy<double>, grid1(m), dim3(256), 0, st, (double*)y,
                           (const double*)x, m);
    else
        launch_kernel(k_copy<float>, grid1(m), dim3(256), 0, st, (float*)y,
                           (const float*)x, m);
    return launch_error();
}

hipError_t launch_cast(void* y, int dty, const void* x, int dtx, long long n, hipStream_t st)
{
    if (dtype_complex(dty) != dtype_complex(dtx)) return hipErrorInvalidValue;
    i64 m = (i64)n * (dtype_complex(dty) ? 2 : 1);
    if (m <= 0) return hipSuccess;
    const bool yd = (dty == DT_F64 || dty == DT_C128), xd = (dtx == DT_F64 || dtx == DT_C128);
    if (yd && !xd)
        launch_kernel((k_cast<double, float>), grid1(m), dim3(256), 0, st, (double*)y,
                           (const float*)x, m);
    else if (!yd && xd)
        launch_kernel((k_cast<float, double>), grid1(m), dim3(256), 0, st, (float*)y,
                           (const double*)x, m);
    else
        return launch_copy(y, x, dty, n, st);
    return launch_error();
}

hipError_t launch_add_potential(void* out, const void* in, const double* v, long long nin,
                                long long ng, int dtype, hipStream_t st)
{
    const i64 n = (i64)nin * ng;
    if (n <= 0) return hipSuccess;
    i64 b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    const dim3 g((unsigned)b);
    switch (dtype) {
    case DT_F64: launch_kernel(k_add_pot<double>, g, dim3(256), 0, st, (double*)out, (const double*)in, v, (i64)ng, n); break;
    case DT_C128: launch_kernel(k_add_pot<double2>, g, dim3(256), 0, st, (double2*)out, (const double2*)in, v, (i64)ng, n); break;
    default: return hipErrorInvalidValue;
    }
    return launch_error();
}

hipError_t launch_scale(double* y, double a, long long n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    launch_kernel(k_scale, grid1(n), dim3(256), 0, st, y, a, (i64)n);
    return launch_error();
}

hipError_t launch_copy_bytes(void* y, const void* x, long long nbytes, hipStream_t st)
{
    const i64 n = nbytes / 16;
    const i64 flat = (n + 255) / 256;
    if (flat >= 1 && flat < 0x7fffffffLL) {
        // one 16-byte element per thread over the whole range: the fastest
        // copy measured on MI355X (every wave's loads in flight at once)
        launch_kernel(k_copy16_flat, dim3((unsigned)flat), dim3(256), 0, st, (uint4*)y,
                      (const uint4*)x, n);
        return launch_error();
    }
    i64 blocks = (n + 1023) / 1024;
    if (blocks > 8192) blocks = 8192;   // 32 blocks per CU, then grid-stride
    if (blocks < 1) blocks = 1;
    launch_kernel(k_copy16, dim3((unsigned)blocks), dim3(256), 0, st, (uint4*)y, (const uint4*)x, n);
    return launch_error();
}

}  // namespace gpaw
