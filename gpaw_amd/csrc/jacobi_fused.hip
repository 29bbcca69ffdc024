// Weighted-Jacobi sweeps fused by temporal blocking, for the small levels
// of the multigrid hierarchy (gfx950).
//
// On a 47^3 or 23^3 level one sweep is a few microseconds of work and about
// as long again of launch and drain, so the coarse levels' eight sweeps per
// V-cycle are dominated by launch boundaries.  Here one launch runs S
// sweeps: a workgroup owns an 8^3 box of the output, loads the box with an
// S-cell halo (16^3 cells for S = 4) into LDS, and sweeps S times between
// two LDS buffers; sweep s updates the cells that sweep s + 1 will read (the
// box plus S - s halo cells), so after S sweeps the box holds exactly the
// values S separate sweeps give.  Halo cells are recomputed by every box
// that needs them, which costs arithmetic on these small grids, not
// bandwidth.
//
// Each cell update is the single-sweep kernel's expression (7-point star,
// c/bmgs/relax.c:47-71 via st_epilogue<ST_JACOBI>: x summed over the six
// neighbours in the order z+, z-, y+, y-, x+, x-, then
// (1-w) f + w (s - x) / c_0), so results are bit-identical.  Cells outside a
// zero-boundary axis stay zero; periodic axes wrap (any number of periods:
// the halo may exceed a small level's extent).  Real data only (the
// Poisson hierarchy); phases are 1.
//
// Optionally the first launch of a level's post-smoothing also applies the
// coarse-grid correction phi -= I e (K = 2 interpolation of the coarser
// level's result, gpaw/poisson.py:266-267) to every cell it loads, with
// interp2_combine's arithmetic: the same values as the interpolate-and-
// subtract launch it replaces.  The block first stages the coarse values
// its cells interpolate from (at most 10 per axis) in LDS.  Cells are
// interpolated at their unwrapped coordinates: on a periodic axis a shift
// by the fine period shifts the coarse index by the coarse period, so the
// value equals the one at the wrapped point.
#include <cstdlib>

#include "kernels.h"
#include "transfer_common.h"

namespace gpaw {

TrArgs make_trargs(const Transfer& tr);

namespace {

// A launch runs S = 4 or 2 sweeps on an E^3 LDS box (E = 16 or 12) whose
// output is the inner B^3 = (E - 2S)^3 cells: 8^3 outputs for four sweeps
// on 16^3, 12^3 for two.  Two-sweep launches recompute fewer halo cells per
// output (4472 cell updates per 1728 outputs against 5984 per 512) at one
// more launch per four sweeps: the better trade on levels with enough 12^3
// boxes to fill the chip.  12^3 boxes (4^3 outputs for four sweeps) serve
// levels too small for enough 8^3 boxes: more recomputed halo per output,
// but eight times the blocks.
constexpr int jf_box(int S, int E) { return E - 2 * S; }
// threads of a block: one per (y, z) column of the box and run of PX planes
constexpr int jf_threads(int E) { return E == 16 ? 1024 : 576; }
constexpr int jf_cb(int E) { return E / 2 + 2; }   // staged coarse values per axis

__device__ __forceinline__ int jf_wrap(int j, int n, int per, bool& out)
{
    if (j >= 0 && j < n) return j;
    if (!per) {
        out = true;
        return 0;
    }
    j %= n;
    return j < 0 ? j + n : j;
}

// 16^3 boxes: 1024 threads, one (y, z) column of the box per 256 threads,
// the box's 16 planes split into four runs of four (four waves per SIMD hide
// the LDS and division latencies of the sweeps); 12^3 boxes: 576 threads,
// four runs of three planes.

// index j of a coarse axis of extent n for the staging: wrapped (any number
// of periods) or flagged outside a zero-boundary axis
__device__ __forceinline__ int jf_cwrap(int j, int n, int per, bool& out)
{
    return jf_wrap(j, n, per, out);
}

// weight of each fused sweep
struct JfWeights {
    double w[4];
};

template <class V, int S, int E>
__global__ __launch_bounds__(jf_threads(E)) void k_jacobi_fused(const V* __restrict__ in, V* __restrict__ out,
                                                        const V* __restrict__ rho, Geom g,
                                                        StarCoefs7 c, JfWeights w, int nb1, int nb2,
                                                        const V* __restrict__ corr, TrArgs ct,
                                                        int zero_in)
{
    constexpr int B = jf_box(S, E);   // output box edge
    constexpr int JF_NT = jf_threads(E);
    constexpr int JF_CB = jf_cb(E);
    static_assert(JF_NT % (E * E) == 0 && E % (JF_NT / (E * E)) == 0,
                  "one thread per (y, z) column of the box and plane run");
    constexpr int PX = E * E * E / JF_NT;   // planes per thread
    __shared__ V P[2][E * E * E];
    __shared__ V C[JF_CB * JF_CB * JF_CB];
    const int b = blockIdx.x;
    const int bz = b % nb2, by = (b / nb2) % nb1, bx = b / (nb2 * nb1);
    const int o0 = bx * B - S, o1 = by * B - S, o2 = bz * B - S;
    const int tz = threadIdx.x % E, ty = (threadIdx.x / E) % E;
    const int x0 = (threadIdx.x / (E * E)) * PX;   // first plane of this thread
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    bool out12 = false;
    const int j1 = jf_wrap(o1 + ty, n1, g.per[1], out12);
    const int j2 = jf_wrap(o2 + tz, n2, g.per[2], out12);
    const int col = j1 * n2 + j2;
    const long plane = (long)n1 * n2;
    typedef typename VT<V>::R RT;
    // every global load is issued before the first is used: this thread's
    // box cells (zero outside a zero-boundary axis) and their rho, then the
    // coarse values of the correction
    V r[PX], v[PX];
    unsigned outside = out12 ? ~0u : 0u;   // bit p: cell (x0 + p, ty, tz) is outside
#pragma unroll
    for (int p = 0; p < PX; p++) {
        bool o = out12;
        const int j0 = jf_wrap(o0 + x0 + p, n0, g.per[0], o);
        if (o) outside |= 1u << p;
        const long idx = (long)j0 * plane + col;
        v[p] = zero_in ? V(0) : in[o ? 0 : idx];
        r[p] = rho[o ? 0 : idx];
    }
    // the coarse values: unwrapped coarse indices from cm
    int cm0 = 0, cm1 = 0, cm2 = 0;
    if (corr) {
        constexpr int NCS = JF_CB * JF_CB * JF_CB;
        constexpr int NC = (NCS + JF_NT - 1) / JF_NT;
        cm0 = ((o0 + ct.skip[0][0]) >> 1) - ct.lo[0];
        cm1 = ((o1 + ct.skip[1][0]) >> 1) - ct.lo[1];
        cm2 = ((o2 + ct.skip[2][0]) >> 1) - ct.lo[2];
        V cv[NC];
        unsigned cmask = 0;   // bit k: staged value k of this thread is a zero ghost
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const int i = threadIdx.x + k * JF_NT;
            cv[k] = V(0);
            if (i < NCS) {
                bool o = false;
                const int c0 = jf_cwrap(cm0 + i / (JF_CB * JF_CB), ct.gin.n[0], ct.gin.per[0], o);
                const int c1 = jf_cwrap(cm1 + (i / JF_CB) % JF_CB, ct.gin.n[1], ct.gin.per[1], o);
                const int c2 = jf_cwrap(cm2 + i % JF_CB, ct.gin.n[2], ct.gin.per[2], o);
                cv[k] = corr[o ? 0 : ((long)c0 * ct.gin.n[1] + c1) * ct.gin.n[2] + c2];
                if (o) cmask |= 1u << k;
            }
        }
#pragma unroll
        for (int k = 0; k < NC; k++)
            if (threadIdx.x + k * JF_NT < NCS)
                C[threadIdx.x + k * JF_NT] = ((cmask >> k) & 1) ? V(0) : cv[k];
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < PX; p++) {
        const bool o = (outside >> p) & 1;
        if (corr && !o) {
            const int q0 = o0 + x0 + p + ct.skip[0][0], q1 = o1 + ty + ct.skip[1][0],
                      q2 = o2 + tz + ct.skip[2][0];
            const int l0 = (q0 >> 1) - ct.lo[0] - cm0, l1 = (q1 >> 1) - ct.lo[1] - cm1,
                      l2 = (q2 >> 1) - ct.lo[2] - cm2;
            v[p] = vsub(v[p], interp2_combine<V>(
                                  [&](int d0, int d1, int d2) {
                                      return C[((l0 + d0) * JF_CB + l1 + d1) * JF_CB + l2 + d2];
                                  },
                                  q0 & 1, q1 & 1, q2 & 1));
        }
        P[0][((x0 + p) * E + ty) * E + tz] = o ? V(0) : v[p];
    }
    __syncthreads();
#pragma unroll
    for (int s = 1; s <= S; s++) {
        const V* __restrict__ src = P[(s - 1) & 1];
        V* __restrict__ dst = P[s & 1];
        if (ty >= s && ty < E - s && tz >= s && tz < E - s) {
#pragma unroll
            for (int p = 0; p < PX; p++) {
                const int x = x0 + p;
                if (x < s || x >= E - s) continue;
                const int cell = (x * E + ty) * E + tz;
                V xs = V(0);
                xs = vmac(xs, src[cell + 1], (RT)c.c[1]);
                xs = vmac(xs, src[cell - 1], (RT)c.c[2]);
                xs = vmac(xs, src[cell + E], (RT)c.c[3]);
                xs = vmac(xs, src[cell - E], (RT)c.c[4]);
                xs = vmac(xs, src[cell + E * E], (RT)c.c[5]);
                xs = vmac(xs, src[cell - E * E], (RT)c.c[6]);
                const V f = st_epilogue<ST_JACOBI>(xs, src[cell], r[p], (RT)w.w[s - 1], (RT)c.c[0]);
                dst[cell] = ((outside >> p) & 1) ? V(0) : f;
            }
        }
        __syncthreads();
    }
    // the box: cells S .. S + B - 1 along each axis, inside the grid
    const int i1 = o1 + ty, i2 = o2 + tz;
    if (ty >= S && ty < S + B && tz >= S && tz < S + B && i1 < n1 && i2 < n2) {
#pragma unroll
        for (int p = 0; p < PX; p++) {
            const int x = x0 + p;
            const int i0 = o0 + x;
            if (x >= S && x < S + B && i0 < n0)
                out[(long)i0 * plane + (long)i1 * n2 + i2] = P[S & 1][(x * E + ty) * E + tz];
        }
    }
}

}  // namespace

bool jacobi_fused_ok(const Stencil& s, const Geom& g)
{
    return s.star_r == 1 && g.n[0] >= 1 && g.n[1] >= 1 && g.n[2] >= 1;
}

int jacobi_fused_edge(const Geom& g, int sweeps)
{
    static const int forced = [] {
        const char* e = getenv("GPAW_AMD_JF_BOX");
        return e ? atoi(e) : 0;
    }();
    if (forced == 16 || forced == 12) return forced;
    // 16^3 boxes down to 64 of them (39^3 with four sweeps: 125); below,
    // 12^3 boxes (23^3: 27 boxes of 16^3 against 216 of 12^3).  Measured
    // both ways on 47^3, 39^3 and 23^3 levels.
    const int B = jf_box(sweeps, 16);
    const long nb = (long)((g.n[0] + B - 1) / B) * ((g.n[1] + B - 1) / B) * ((g.n[2] + B - 1) / B);
    return nb >= 64 ? 16 : 12;
}

int jacobi_fused_sweeps(const Geom& g)
{
    // two sweeps per launch once the 12^3 boxes fill the chip (256 CUs)
    const long nb = (long)((g.n[0] + 11) / 12) * ((g.n[1] + 11) / 12) * ((g.n[2] + 11) / 12);
    return nb >= 256 ? 2 : JF_SWEEPS;
}

hipError_t launch_jacobi_fused(const Stencil& s, int dtype, const void* in, void* out,
                               const void* rho, const Geom& g, double w, hipStream_t st,
                               const void* corr, const Transfer* ctr, bool zero_in, int sweeps,
                               const double* ws)
{
    if (!jacobi_fused_ok(s, g) || (sweeps != 2 && sweeps != 4)) return hipErrorInvalidValue;
    TrArgs ct{};
    if (corr) {
        // the correction is a K = 2 interpolation onto exactly this grid
        if (!ctr || ctr->K != 2 || !ctr->interpolate) return hipErrorInvalidValue;
        ct = make_trargs(*ctr);
        for (int a = 0; a < 3; a++)
            if (ct.nout[a] != g.n[a]) return hipErrorInvalidValue;
    }
    StarCoefs7 c;
    for (int k = 0; k < 7; k++) c.c[k] = s.coefs[k];
    JfWeights jw;
    for (int k = 0; k < 4; k++) jw.w[k] = ws && k < sweeps ? ws[k] : w;
    const int E = jacobi_fused_edge(g, sweeps);
    const int B = jf_box(sweeps, E);
    const int nb0 = (g.n[0] + B - 1) / B, nb1 = (g.n[1] + B - 1) / B, nb2 = (g.n[2] + B - 1) / B;
    const dim3 grid((unsigned)(nb0 * nb1 * nb2));
#define GPAW_JF(T, S, EE)                                                                    \
    launch_kernel(k_jacobi_fused<T, S, EE>, grid, dim3(jf_threads(EE)), 0, st, (const T*)in,  \
                  (T*)out, (const T*)rho, g, c, jw, nb1, nb2, (const T*)corr, ct, (int)zero_in)
#define GPAW_JF_E(T, S)                                                                      \
    if (E == 16)                                                                             \
        GPAW_JF(T, S, 16);                                                                   \
    else                                                                                     \
        GPAW_JF(T, S, 12);
    if (dtype == DT_F64) {
        if (sweeps == 4) {
            GPAW_JF_E(double, 4)
        } else {
            GPAW_JF_E(double, 2)
        }
    } else if (dtype == DT_F32) {
        if (sweeps == 4) {
            GPAW_JF_E(float, 4)
        } else {
            GPAW_JF_E(float, 2)
        }
    } else {
        return hipErrorInvalidValue;
    }
#undef GPAW_JF_E
#undef GPAW_JF
    return launch_error();
}

}  // namespace gpaw
