// Fused coarse-level V-cycle (gfx950): the smallest levels of the multigrid
// hierarchy are smoothed, restricted and corrected by ONE workgroup of 1024
// threads in one launch, with a workgroup barrier between passes instead of
// a kernel boundary.  On these grids (11^3 and below for a 95^3 problem) a
// separate launch per pass costs more than the pass itself.  When the level
// arrays fit, they live in LDS: only the first level's right-hand side is
// read from HBM and its solution written back.
//
// The passes and their arithmetic are those of iterate2
// (gpaw/poisson.py:245-270) on 7-point operators (Laplace nn=1) with K=2
// transformers, exactly as the standalone kernels evaluate them:
//   Jacobi    x = sum_{c>=1} coef_c phi[g+off_c] (order z+, z-, y+, y-, x+, x-),
//             phi' = (1-w) phi + w (rho - x) / coef_0   (c/bmgs/relax.c:47-71)
//   residual  r = sum_c coef_c phi[g+off_c] - rho       (c/bmgs/fd.c:41-44)
//   restrict / interpolate   transfer_common.h
// so the fused path is bit-identical to launching the passes one by one.
#include "kernels.h"
#include "coarse.h"
#include "transfer_common.h"

namespace gpaw {

namespace {

__device__ __forceinline__ void split(int idx, const Geom& g, int& i0, int& i1, int& i2)
{
    i2 = idx % g.n[2];
    const int t = idx / g.n[2];
    i1 = t % g.n[1];
    i0 = t / g.n[1];
}

// level arrays: phi A/B, rho, residual
enum { PA = 0, PB = 1, RHO = 2, RES = 3 };

// A level has at most 2048 points (two per thread).  Each thread's points
// and their six neighbours (z+, z-, y+, y-, x+, x-: the summation order of
// the 7-point operator, centre first) come from the plan's table; a
// neighbour outside a zero boundary is -1 in the table.  ZS (the LDS
// layout): such a neighbour reads the level's zero slot instead, element n
// of its phi arrays, which holds 0.0 for the whole kernel.
struct Pts {
    int idx[2];       // flat index, -1 when the thread has no such point
    int nb[2][6];
};

template <bool ZS>
__device__ __forceinline__ void level_points(const CoarseLevel& L, const short* __restrict__ tab,
                                             Pts& p)
{
    const int n = L.g.n[0] * L.g.n[1] * L.g.n[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int idx = threadIdx.x + 1024 * k;
        p.idx[k] = idx < n ? idx : -1;
        const short* t = tab + L.nb_off + 6 * (idx < n ? idx : 0);
#pragma unroll
        for (int m = 0; m < 6; m++) p.nb[k][m] = (ZS && t[m] < 0) ? n : t[m];
    }
}

// x + the six neighbour terms in summation order (the coarse levels are
// real, phases 1, so a neighbour is a plain load or zero).  A zero
// neighbour contributes 0.0 * c either way: the same bits with or without
// the zero slot.
template <bool ZS>
__device__ __forceinline__ double nb_sum(const double* __restrict__ a, const int* nb,
                                         const double* c, double x)
{
    // every load issued unconditionally (without the zero slot: index 0 for
    // a zero neighbour, whose value is then replaced by 0): no branch
    // between the loads
    double v[6];
#pragma unroll
    for (int m = 0; m < 6; m++) v[m] = a[(ZS || nb[m] >= 0) ? nb[m] : 0];
#pragma unroll
    for (int m = 0; m < 6; m++) x = x + ((ZS || nb[m] >= 0) ? v[m] : 0.0) * c[m + 1];
    return x;
}

// The four arrays of one level.  In LDS mode they are derived from the
// kernel's shared array, so the compiler addresses them as LDS (ds_read /
// ds_write); from a pointer table in memory it could only use flat
// accesses, which take the vector-memory path even for LDS.
struct LevelArrays {
    double* a[4];
};

template <bool LDS>
__device__ __forceinline__ LevelArrays level_arrays(const CoarseLevel& L, double* smem)
{
    LevelArrays r;
    if (LDS) {   // n + 1 elements each: element n of phi A / B is the zero slot
        const int n = L.g.n[0] * L.g.n[1] * L.g.n[2];
#pragma unroll
        for (int k = 0; k < 4; k++) r.a[k] = smem + L.off + k * (n + 1);
    } else {
        r.a[PA] = L.phiA;
        r.a[PB] = L.phiB;
        r.a[RHO] = L.rho;
        r.a[RES] = L.res;
    }
    return r;
}

#ifndef GPAW_COARSE_DIV_RN
#define GPAW_COARSE_DIV_RN 1
#endif

// a / d rounded to nearest, from r = 1 / d rounded to nearest: the product
// q = a r is within an ulp of a / d, its remainder a - q d is exact in one
// fma, and q + (a - q d) r rounded once is the correctly rounded quotient
// (Markstein's theorem, binary64 without overflow or underflow).  A zero
// remainder keeps q, so exact quotients and signed zeros come out as a / d
// does.  Three operations instead of the division sequence, on a pass that
// runs on one compute unit.
__device__ __forceinline__ double div_rn(double a, double d, double r)
{
    const double q = a * r;
    const double e = __builtin_fma(-q, d, a);
    return e == 0.0 ? q : __builtin_fma(e, r, q);
}

template <bool ZS>
__device__ __forceinline__ void jacobi_sweeps(const CoarseLevel& L, const Pts& p,
                                              const LevelArrays& A, int k, int& cur)
{
    const double* __restrict__ rho = A.a[RHO];
    double c[7];
#pragma unroll
    for (int m = 0; m < 7; m++) c[m] = L.c[m];
    const double w = L.w;
    const double rc0 = 1.0 / c[0];
    for (int it = 0; it < k; it++) {
        const double* __restrict__ src = cur == PA ? A.a[PA] : A.a[PB];
        double* __restrict__ dst = cur == PA ? A.a[PB] : A.a[PA];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int idx = p.idx[q];
            if (idx < 0) continue;
            const double x = nb_sum<ZS>(src, p.nb[q], c, 0.0);
            if (GPAW_COARSE_DIV_RN)
                dst[idx] = (1.0 - w) * src[idx] + div_rn(w * (rho[idx] - x), c[0], rc0);
            else
                dst[idx] = (1.0 - w) * src[idx] + w * (rho[idx] - x) / c[0];
        }
        __syncthreads();
        cur ^= 1;
    }
}

}  // namespace

// TB: storage type of the first level's right-hand side and solution in HBM
// (double, or float when the fused kernel is the FP64 coarse solve of an
// FP32 V-cycle; then the arrays must live in LDS and are converted there).
// LDS: the level arrays live in the kernel's shared memory (plan->lds).
template <class TB, bool LDS>
__global__ __launch_bounds__(1024) void k_coarse_vcycle(const CoarsePlan* __restrict__ Pg)
{
    extern __shared__ double smem[];
    // the plan is read from an LDS copy: its fields are read at every phase,
    // and from global memory each such read would miss the scalar cache
    // (the kernel runs once per V-cycle, between much larger kernels)
    __shared__ CoarsePlan Ps;
    {
        static_assert(sizeof(CoarsePlan) % 4 == 0, "plan copied in 4-byte words");
        const int* src = (const int*)Pg;
        int* dst = (int*)&Ps;
        for (int i = threadIdx.x; i < (int)(sizeof(CoarsePlan) / 4); i += blockDim.x)
            dst[i] = src[i];
        __syncthreads();
    }
    const CoarsePlan* P = &Ps;
    // neighbour tables: copied next to the level arrays in LDS mode
    const short* tab = P->nbtab;
    if (LDS) {
        int* dst = (int*)(smem + P->lds_doubles);
        const int* src = (const int*)P->nbtab;
        for (int i = threadIdx.x; i < P->nb_words; i += blockDim.x) dst[i] = src[i];
        tab = (const short*)dst;
        // the zero slots (LDS starts undefined)
        for (int l = threadIdx.x; l < P->nlev; l += blockDim.x) {
            const CoarseLevel& L = P->lev[l];
            const int n = L.g.n[0] * L.g.n[1] * L.g.n[2];
            smem[L.off + n] = 0.0;
            smem[L.off + (n + 1) + n] = 0.0;
        }
        // made visible by the barrier after the first level's load below
    }
    Phases ph;
    for (int a = 0; a < 3; a++)
        for (int d = 0; d < 2; d++) {
            ph.re[a][d] = 1.0;
            ph.im[a][d] = 0.0;
        }
    const int nl = P->nlev;
    if (LDS) {   // right-hand side of the first coarse level, phi = 0
        const CoarseLevel& G = P->lev[0];
        const LevelArrays A = level_arrays<LDS>(G, smem);
        const int n = G.g.n[0] * G.g.n[1] * G.g.n[2];
        const TB* __restrict__ grho = (const TB*)G.rho;
        for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
            A.a[RHO][idx] = (double)grho[idx];
            A.a[PA][idx] = 0.0;
        }
        __syncthreads();
    }
    // down: pre-smooth, residual, restrict (phi of the next level zeroed)
    for (int l = 0; l < nl - 1; l++) {
        const CoarseLevel& L = P->lev[l];
        const LevelArrays A = level_arrays<LDS>(L, smem);
        Pts pts;
        level_points<LDS>(L, tab, pts);
        int cur = PA;
        jacobi_sweeps<LDS>(L, pts, A, L.pre, cur);
        const double* __restrict__ phi = cur == PA ? A.a[PA] : A.a[PB];
        const double* __restrict__ rho = A.a[RHO];
        double* __restrict__ res = A.a[RES];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int idx = pts.idx[q];
            if (idx < 0) continue;
            double c[7];
#pragma unroll
            for (int m = 0; m < 7; m++) c[m] = L.c[m];
            res[idx] = nb_sum<LDS>(phi, pts.nb[q], c, 0.0 + phi[idx] * c[0]) - rho[idx];
        }
        __syncthreads();
        const CoarseLevel& N = P->lev[l + 1];
        const LevelArrays NA = level_arrays<LDS>(N, smem);
        const int m = N.g.n[0] * N.g.n[1] * N.g.n[2];
        for (int idx = threadIdx.x; idx < m; idx += blockDim.x) {
            int o0, o1, o2;
            split(idx, N.g, o0, o1, o2);
            NA.a[RHO][idx] = restrict_point<double, 2>(res, o0, o1, o2, L.rst, ph);
            NA.a[PA][idx] = 0.0;
        }
        __syncthreads();
    }
    {   // bottom: post-smooth only
        const CoarseLevel& L = P->lev[nl - 1];
        const LevelArrays A = level_arrays<LDS>(L, smem);
        Pts pts;
        level_points<LDS>(L, tab, pts);
        int cur = PA;
        jacobi_sweeps<LDS>(L, pts, A, L.post, cur);
    }
    // up: correct from the coarser level, post-smooth
    for (int l = nl - 2; l >= 0; l--) {
        const CoarseLevel& L = P->lev[l];
        const CoarseLevel& N = P->lev[l + 1];
        const LevelArrays A = level_arrays<LDS>(L, smem);
        const LevelArrays NA = level_arrays<LDS>(N, smem);
        int cur = (L.pre & 1) ? PB : PA;
        const int kc = (l + 1 < nl - 1 ? N.pre : 0) + N.post;
        const double* __restrict__ coarse = (kc & 1) ? NA.a[PB] : NA.a[PA];
        double* __restrict__ phi = cur == PA ? A.a[PA] : A.a[PB];
        // one coarse cell (2 x 2 x 2 fine points) per thread, as k_interp2
        const int nc0 = interp2_cells(L.itp, 0), nc1 = interp2_cells(L.itp, 1),
                  nc2 = interp2_cells(L.itp, 2);
        for (int idx = threadIdx.x; idx < nc0 * nc1 * nc2; idx += blockDim.x) {
            const int k2 = idx % nc2, k1 = (idx / nc2) % nc1, k0 = idx / (nc2 * nc1);
            interp2_cell<double, TR_SUBTRACT>(coarse, phi, L.itp, k0, k1, k2, ph);
        }
        __syncthreads();
        Pts pts;
        level_points<LDS>(L, tab, pts);
        jacobi_sweeps<LDS>(L, pts, A, L.post, cur);
    }
    if (LDS) {   // hand the first level's solution back where the caller expects it
        const CoarseLevel& G = P->lev[0];
        const LevelArrays A = level_arrays<LDS>(G, smem);
        const int k = (nl > 1 ? G.pre : 0) + G.post;
        const double* src = (k & 1) ? A.a[PB] : A.a[PA];
        TB* dst = (TB*)((k & 1) ? G.phiB : G.phiA);
        const int n = G.g.n[0] * G.g.n[1] * G.g.n[2];
        for (int idx = threadIdx.x; idx < n; idx += blockDim.x) dst[idx] = (TB)src[idx];
    }
}

// The single-coarse-level case (the hierarchy's last level alone, as for a
// 159^3 problem, whose 19^3 level is too large for one workgroup and 9^3
// is the last): post-smoothing sweeps only, in a compact kernel whose
// parameters arrive as kernel arguments.  The generic kernel first copies
// its plan and neighbour table in, two dependent round trips; here the
// neighbour indices are computed per point.  Same sweep arithmetic as
// jacobi_sweeps.
namespace {

// flat index of (j0, j1, j2), wrapped on periodic axes, -1 outside a zero
// boundary (the neighbour table's rule)
__device__ __forceinline__ int one_nb(int j0, int j1, int j2, const Geom& g)
{
    const int j[3] = {j0, j1, j2};
    int w[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        w[a] = j[a];
        if (w[a] < 0 || w[a] >= g.n[a]) {
            if (!g.per[a]) return -1;
            w[a] = w[a] < 0 ? w[a] + g.n[a] : w[a] - g.n[a];
        }
    }
    return (w[0] * g.n[1] + w[1]) * g.n[2] + w[2];
}

template <class TB>
__global__ __launch_bounds__(1024) void k_coarse_one(CoarseOne p)
{
    extern __shared__ double smem[];
    const int n = p.g.n[0] * p.g.n[1] * p.g.n[2];
    // phi A / B with their zero slots (element n), then rho
    double* A = smem;
    double* B = smem + (n + 1);
    double* rho = smem + 2 * (n + 1);
    if (threadIdx.x == 0) {
        A[n] = 0.0;
        B[n] = 0.0;
    }
    Pts q;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int idx = threadIdx.x + 1024 * k;
        q.idx[k] = idx < n ? idx : -1;
        int i0 = 0, i1 = 0, i2 = 0;
        if (idx < n) split(idx, p.g, i0, i1, i2);
        q.nb[k][0] = one_nb(i0, i1, i2 + 1, p.g);
        q.nb[k][1] = one_nb(i0, i1, i2 - 1, p.g);
        q.nb[k][2] = one_nb(i0, i1 + 1, i2, p.g);
        q.nb[k][3] = one_nb(i0, i1 - 1, i2, p.g);
        q.nb[k][4] = one_nb(i0 + 1, i1, i2, p.g);
        q.nb[k][5] = one_nb(i0 - 1, i1, i2, p.g);
#pragma unroll
        for (int m = 0; m < 6; m++)
            if (q.nb[k][m] < 0) q.nb[k][m] = n;
        if (idx < n) {
            rho[idx] = (double)((const TB*)p.rho)[idx];
            A[idx] = 0.0;
        }
    }
    __syncthreads();
    double c[7];
#pragma unroll
    for (int m = 0; m < 7; m++) c[m] = p.c[m];
    const double w = p.w;
    const double rc0 = 1.0 / c[0];
    double* src = A;
    double* dst = B;
    for (int it = 0; it < p.sweeps; it++) {
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int idx = q.idx[k];
            if (idx < 0) continue;
            const double x = nb_sum<true>(src, q.nb[k], c, 0.0);
            if (GPAW_COARSE_DIV_RN)
                dst[idx] = (1.0 - w) * src[idx] + div_rn(w * (rho[idx] - x), c[0], rc0);
            else
                dst[idx] = (1.0 - w) * src[idx] + w * (rho[idx] - x) / c[0];
        }
        __syncthreads();
        double* t = src;
        src = dst;
        dst = t;
    }
#pragma unroll
    for (int k = 0; k < 2; k++)
        if (q.idx[k] >= 0) ((TB*)p.out)[q.idx[k]] = (TB)src[q.idx[k]];
}

}  // namespace

// the single-level kernel's arguments from a one-level plan
static bool coarse_one_args(const CoarsePlan& h, CoarseOne& a)
{
    if (h.nlev != 1) return false;
    const CoarseLevel& L = h.lev[0];
    const long n = (long)L.g.n[0] * L.g.n[1] * L.g.n[2];
    if (n < 1 || n > 2048) return false;
    a.g = L.g;
    for (int m = 0; m < 7; m++) a.c[m] = L.c[m];
    a.w = L.w;
    a.sweeps = L.post;
    a.rho = L.rho;
    a.out = (L.post & 1) ? L.phiB : L.phiA;
    return true;
}

static int nb_host(int j0, int j1, int j2, const Geom& g)
{
    const int j[3] = {j0, j1, j2};
    int w[3];
    for (int a = 0; a < 3; a++) {
        w[a] = j[a];
        if (w[a] < 0 || w[a] >= g.n[a]) {
            if (!g.per[a]) return -1;
            w[a] = w[a] < 0 ? w[a] + g.n[a] : w[a] - g.n[a];
        }
    }
    return (w[0] * g.n[1] + w[1]) * g.n[2] + w[2];
}

void coarse_neighbour_table(const Geom& g, short* tab)
{
    for (int i0 = 0; i0 < g.n[0]; i0++)
        for (int i1 = 0; i1 < g.n[1]; i1++)
            for (int i2 = 0; i2 < g.n[2]; i2++) {
                short* t = tab + 6 * ((i0 * g.n[1] + i1) * g.n[2] + i2);
                t[0] = (short)nb_host(i0, i1, i2 + 1, g);
                t[1] = (short)nb_host(i0, i1, i2 - 1, g);
                t[2] = (short)nb_host(i0, i1 + 1, i2, g);
                t[3] = (short)nb_host(i0, i1 - 1, i2, g);
                t[4] = (short)nb_host(i0 + 1, i1, i2, g);
                t[5] = (short)nb_host(i0 - 1, i1, i2, g);
            }
}

hipError_t coarse_set_lds_limit(int bytes)
{
    hipError_t e = hipFuncSetAttribute((const void*)k_coarse_vcycle<double, true>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)k_coarse_vcycle<float, true>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    return e;
}

// lds_bytes > 0: the plan keeps its level arrays in LDS (plan->lds)
hipError_t launch_coarse_vcycle(const CoarsePlan* d_plan, const CoarsePlan& h_plan, int lds_bytes,
                                hipStream_t st)
{
    CoarseOne one;
    if (coarse_one_args(h_plan, one)) {
        const int n = one.g.n[0] * one.g.n[1] * one.g.n[2];
        launch_kernel(k_coarse_one<double>, dim3(1), dim3(1024),
                      3 * (n + 1) * (int)sizeof(double), st, one);
        return launch_error();
    }
    if (lds_bytes > 0)
        launch_kernel(k_coarse_vcycle<double, true>, dim3(1), dim3(1024), lds_bytes, st, d_plan);
    else
        launch_kernel(k_coarse_vcycle<double, false>, dim3(1), dim3(1024), 0, st, d_plan);
    return launch_error();
}

hipError_t launch_coarse_vcycle_f32(const CoarsePlan* d_plan, const CoarsePlan& h_plan,
                                    int lds_bytes, hipStream_t st)
{
    CoarseOne one;
    if (coarse_one_args(h_plan, one)) {
        const int n = one.g.n[0] * one.g.n[1] * one.g.n[2];
        launch_kernel(k_coarse_one<float>, dim3(1), dim3(1024),
                      3 * (n + 1) * (int)sizeof(double), st, one);
        return launch_error();
    }
    if (lds_bytes <= 0) return hipErrorInvalidValue;   // float levels need LDS mode
    launch_kernel(k_coarse_vcycle<float, true>, dim3(1), dim3(1024), lds_bytes, st, d_plan);
    return launch_error();
}

}  // namespace gpaw
