// Mixed-precision multigrid Poisson solver (gfx950).
//
// Precision split:
//   FP64  residual d = L phi - rho (the reference operator, c/bmgs/fd.c),
//         correction accumulation into phi, all dot products and the
//         convergence test, the charge-neutrality terms (Python side), and
//         the coarsest levels (the fused LDS coarse kernel computes in FP64;
//         only its first-level right-hand side and solution are FP32);
//   FP32  smoothing (weighted Jacobi), residuals, restriction and
//         interpolation of the inner V-cycle on every level above that.
//
// Outer loop (FP64), either
//   iterative refinement  d = L phi - rho; e = V^k(d); phi -= e
//   flexible PCG          preconditioned conjugate gradients with one inner
//                         V-cycle as preconditioner (Polak-Ribiere beta,
//                         which tolerates the preconditioner's FP32
//                         rounding), i.e. iterative refinement with a Krylov
//                         recurrence on the corrections; when the recursive
//                         |d| meets the target the true residual is
//                         recomputed, and the iteration restarts from it if
//                         it does not.
// The inner V-cycle uses the reference hierarchy (gpaw/poisson.py:64-113)
// with symmetric level-0 smoothing (2 pre + 2 post Chebyshev-weighted Jacobi
// sweeps by default, mixed_schedule) so that it is
// a symmetric preconditioner.  It runs in FP32, or -- to measure what the
// precision itself contributes -- the very same V-cycle in FP64 inside the
// same outer loop.  Stopping rule: ||d||_2 <= tol * ||rho||_2.
//
// Scalars (alpha, beta, dot products) stay on the device; the host reads
// one 8-byte norm per chunk of outer iterations.  The inner V-cycle and the
// per-iteration kernel sequences are captured as hipGraphs.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "poisson_internal.h"

using namespace gpaw;

// One inner V-cycle hierarchy: its arrays and storage type.  The FP32
// hierarchy owns every level; the FP64 one (used to attribute the speed-up:
// the same outer loop with the inner V-cycle in FP64) owns level 0 and
// reuses the plan's own arrays below it, which are scratch during a solve.
struct VLevels {
    int dt = DT_F32;
    std::vector<void*> phiA, phiB, rho, res;
    std::vector<void*> owned;
    CoarsePlan* d_coarse = nullptr;     // fused coarse levels (FP64 inside)
    CoarsePlan h_coarse{};              // its host copy
    bool own_coarse = false;
    int lc = -1;                        // first fused coarse level (-1: none)
};

struct MixedState {
    int L = 0;
    std::vector<int> pre, post;
    // level-0 Jacobi weights per sweep (empty: the plan's weight for every
    // sweep): pre-smoothing in order, post-smoothing in the reverse order
    std::vector<double> w0pre, w0post;
    // the same for levels 1 .. (empty: the plan's weight)
    std::vector<std::vector<double>> wpre, wpost;
    VLevels v[2];                       // [0] FP32, [1] FP64 inner V-cycle
    bool ready[2] = {false, false};
    double *p = nullptr, *q = nullptr;  // FP64 search direction, L p
    double* S = nullptr;                // device scalars
    double* h_S = nullptr;              // pinned readback
    double* partial = nullptr;
    // graphs per inner precision
    hipGraphExec_t g_vcycle[2] = {nullptr, nullptr};   // one V-cycle from zero
    hipGraphExec_t g_vwarm[2] = {nullptr, nullptr};    // one V-cycle from the last result
    hipGraphExec_t g_a[2] = {nullptr, nullptr};        // PCG: q = Lp, alpha, update, |d|^2
    hipGraphExec_t g_b[2] = {nullptr, nullptr};        // PCG: z = V(d), beta, p = z + beta p
    // k PCG steps (b then a) per launch, k = 1 .. MIXED_CHUNK_MAX; [1]: the
    // same followed by the true residual's norm (chunks expected to converge)
    std::vector<hipGraphExec_t> g_ba[2][2];
    // PCG steps the last converged solve took, per inner precision, and the
    // tolerance it was solved to: the first chunk of the next solve to the
    // same tolerance
    int hint[2] = {0, 0};
    double hint_tol[2] = {0.0, 0.0};
    int tuned[2] = {0, 0};              // launch shapes tuned (tuner on)
};

#define MIXED_CHUNK_MAX 16
#define MIXED_PREDICT_MAX 8   // steps per rate-predicted chunk
#define MIXED_MAX_SWEEPS 16   // level-0 sweeps per smoothing phase

// device scalars
// S_DZ = 0 marks a (re)start: the next direction update takes p = z.
// S_D0: |d|^2 of the residual a solve starts from; S_DT: |L phi - rho|^2
// after a chunk that ends with the true residual.
// S_HIST + i: |d|^2 after step i of a chunk (all but its last step; the
// last one's is S_DD).
enum {
    S_DZ = 0, S_PQ = 1, S_DD = 2, S_D0 = 3, S_RR = 4, S_ALPHA = 5, S_DT = 6, S_HIST = 8,
    S_N = S_HIST + MIXED_CHUNK_MAX
};

namespace {

constexpr int NB = 1024;   // blocks of the reduction kernels
// regions of the partials buffer: d.z and z.q (k_dot_dz_zq, [0, 2 NB)),
// |d|^2 of the update, p.q of the separate dot pass
constexpr int PART_DD = 2 * NB, PART_PQ = 3 * NB;

__global__ __launch_bounds__(256) void k_dot(const double* __restrict__ x,
                                             const double* __restrict__ y, i64 n,
                                             double* __restrict__ partial)
{
    __shared__ double red[256];
    double s = 0.0;
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += (i64)gridDim.x * 256)
        s += x[i] * y[i];
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// S[slot] = sum of partial[0..n)
__global__ __launch_bounds__(256) void k_red1(const double* __restrict__ p, int n,
                                              double* __restrict__ S, int slot)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += p[i];
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) S[slot] = s;
}

// Sum of n partials in k_red1's order (every block of a launch gets the
// same value).
__device__ __forceinline__ double sum_partials(const double* __restrict__ p, int n, double* red)
{
    double t = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) t += p[i];
    return block_sum<256>(t, red);
}

// alpha = dz / pq ; phi -= alpha p ; d -= alpha q ; dlow = d (the inner
// V-cycle's right-hand side in its own precision; nullptr when the inner
// V-cycle reads d itself) ; partials of |d|^2 into `partial`.  dz (the ndz
// partials at dzp) and pq (the npq partials at pqp, or S[PQ] when npq = 0)
// are summed by every block itself; block 0 keeps dz in S[DZ] (the next
// step's old dz) and alpha in S[ALPHA].
template <class T>
__global__ __launch_bounds__(256) void k_pcg_update(double* __restrict__ phi, double* __restrict__ d,
                                                    const double* __restrict__ p,
                                                    const double* __restrict__ q,
                                                    T* __restrict__ dlow, double* __restrict__ S,
                                                    i64 n, double* __restrict__ partial,
                                                    const double* __restrict__ pqp, int npq,
                                                    const double* __restrict__ dzp, int ndz)
{
    __shared__ double red[256];
    const double pq = npq > 0 ? sum_partials(pqp, npq, red) : S[S_PQ];
    const double dz = sum_partials(dzp, ndz, red);
    // pq = 0 only when p = 0 (a zero residual): no step
    const double alpha = pq != 0.0 ? dz / pq : 0.0;
    double s = 0.0;
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += (i64)gridDim.x * 256) {
        phi[i] = phi[i] - alpha * p[i];
        const double di = d[i] - alpha * q[i];
        d[i] = di;
        if (dlow) dlow[i] = (T)di;
        s += di * di;
    }
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S[S_ALPHA] = alpha;
        S[S_DZ] = dz;
    }
}

// partials of d.z and z.q (z from the inner V-cycle)
template <class T>
__global__ __launch_bounds__(256) void k_dot_dz_zq(const double* __restrict__ d,
                                                   const T* __restrict__ z,
                                                   const double* __restrict__ q, i64 n,
                                                   double* __restrict__ partial)
{
    __shared__ double red[256];
    double a = 0.0, b = 0.0;
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += (i64)gridDim.x * 256) {
        const double zi = (double)z[i];
        a += d[i] * zi;
        b += zi * q[i];
    }
    a = block_sum<256>(a, red);
    b = block_sum<256>(b, red);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = a;
        partial[gridDim.x + blockIdx.x] = b;
    }
}

// p = z + beta p, beta = -alpha (z.q) / (d_old.z_old) (Polak-Ribiere), z.q
// summed by every block from the nzq partials at zqp, d_old.z_old = S[DZ]
// (kept by the previous update); first, or S[DZ] = 0 (a restart): p = z.
// hist: block 0 also sums the previous step's |d|^2 partials (ndd at ddp)
// into *hist, so the host sees |d| after every step of a chunk.
template <class T>
__global__ __launch_bounds__(256) void k_p_update(double* __restrict__ p, const T* __restrict__ z,
                                                  double* __restrict__ S, i64 n, int first,
                                                  const double* __restrict__ zqp, int nzq,
                                                  const double* __restrict__ ddp, int ndd,
                                                  double* __restrict__ hist)
{
    __shared__ double red[256];
    if (hist && blockIdx.x == 0) {
        const double v = sum_partials(ddp, ndd, red);
        if (threadIdx.x == 0) *hist = v;
    }
    const double dzo = S[S_DZ];
    const bool f = first || dzo == 0.0;   // uniform over the grid
    const double zq = f ? 0.0 : sum_partials(zqp, nzq, red);
    const double beta = f ? 0.0 : -S[S_ALPHA] * zq / dzo;
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += (i64)gridDim.x * 256)
        p[i] = f ? (double)z[i] : (double)z[i] + beta * p[i];
}

// The first Jacobi sweep from phi = 0: every stencil term is zero, so the
// sweep is element-wise: the tile kernel's epilogue with a zero sum and a
// zero centre value (same operations, same bits).
template <class T>
__global__ __launch_bounds__(256) void k_jacobi_from_zero(T* __restrict__ out,
                                                          const T* __restrict__ rho, i64 n,
                                                          double w, double c0)
{
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += (i64)gridDim.x * 256)
        out[i] = st_epilogue<ST_JACOBI>(T(0), T(0), rho[i], (T)w, (T)c0);
}

// phi -= e (FP64 accumulation of the inner V-cycle's correction)
template <class T>
__global__ __launch_bounds__(256) void k_sub_corr(double* __restrict__ phi, const T* __restrict__ e,
                                                  i64 n)
{
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += (i64)gridDim.x * 256)
        phi[i] = phi[i] - (double)e[i];
}

dim3 gridn(i64 n)
{
    i64 b = (n + 255) / 256;
    return dim3((unsigned)(b < NB ? b : NB));
}

// `corr`: the first fused launch first applies phi -= I corr (the
// coarse-grid correction, same values as the separate launch)
hipError_t relax_v(gpaw_poisson* ps, VLevels& V, int l, int k, void*& cur, void*& alt,
                   hipStream_t st, const void* corr = nullptr, bool zero_in = false,
                   const double* wts = nullptr)
{
    Phases ph;
    gpaw_phases(nullptr, ph);
    gpaw_operator* op = ps->ops[l];
    // below level 0, runs of 4 (2 on large levels) sweeps are one fused launch (callers
    // follow `cur`, not the sweep parity)
    if (l > 0 && ps->fuse_jacobi && jacobi_fused_ok(op->st, op->g)) {
        int S = jacobi_fused_sweeps(op->g);
        if (k < S && k >= 2) S = 2;   // runs of two (schedules with fewer sweeps)
        for (; k >= S; k -= S) {
            hipError_t e = launch_jacobi_fused(op->st, V.dt, cur, alt, V.rho[l], op->g, ps->w[l],
                                               st, corr, corr ? &ps->itp[l]->t : nullptr,
                                               zero_in, S, wts);
            if (wts) wts += S;
            corr = nullptr;
            zero_in = false;
            if (e != hipSuccess) return e;
            std::swap(cur, alt);
        }
    }
    for (int it = 0; it < k; it++) {
        const double w = wts ? wts[it] : ps->w[l];
        hipError_t e = corr ? launch_jacobi_corr(op->st, V.dt, cur, alt, V.rho[l], corr, op->g, w,
                                                 ps->itp[l]->t, st)
                            : launch_stencil(op->st, V.dt, ST_JACOBI, cur, alt, V.rho[l], op->g,
                                             ph, 1, w, nullptr, nullptr, true, st);
        corr = nullptr;
        if (e != hipSuccess) return e;
        std::swap(cur, alt);
    }
    return hipSuccess;
}

// one inner V-cycle on level l (level 0: rho = the current residual, result
// in phiA[0]); `res` returns the buffer that holds level l's result
hipError_t emit_v(gpaw_poisson* ps, MixedState* m, VLevels& V, int l, bool zero0, void*& res,
                  hipStream_t st)
{
    Phases ph;
    gpaw_phases(nullptr, ph);
    if (l == V.lc) {   // the fused coarse kernel leaves it where single sweeps would
        const int kc = (l < m->L ? m->pre[l] : 0) + m->post[l];
        res = (kc % 2) ? V.phiB[l] : V.phiA[l];
        return V.dt == DT_F32 ? launch_coarse_vcycle_f32(V.d_coarse, V.h_coarse, ps->coarse_lds, st)
                              : launch_coarse_vcycle(V.d_coarse, V.h_coarse, ps->coarse_lds, st);
    }
    gpaw_operator* op = ps->ops[l];
    void* cur = V.phiA[l];
    void* alt = V.phiB[l];
    hipError_t e = hipSuccess;
    int pre = l < m->L ? m->pre[l] : 0;
    const double* wpre = (l == 0 && !m->w0pre.empty()) ? m->w0pre.data() : nullptr;
    const double* wpost = (l == 0 && !m->w0post.empty()) ? m->w0post.data() : nullptr;
    if (l > 0 && l < (int)m->wpre.size() && !m->wpre[l].empty()) {
        wpre = m->wpre[l].data();
        wpost = m->wpost[l].data();
    }
    if (l == 0 && zero0 && pre >= 2 && ps->fuse_jacobi &&
        jacobi_two_from_zero_ok(op->st, V.dt, op->g)) {
        // the first two sweeps from zero in one launch, into cur (where two
        // single sweeps leave them)
        e = launch_jacobi_two_from_zero(op->st, V.dt, V.rho[0], cur, op->g,
                                        wpre ? wpre[0] : ps->w[0], wpre ? wpre[1] : ps->w[0], st);
        pre -= 2;
        if (wpre) wpre += 2;
    } else if (l == 0 && zero0 && pre > 0 && ps->fuse_jacobi) {
        // the first sweep from zero, element-wise, into alt (as a sweep would)
        if (V.dt == DT_F32)
            launch_kernel(k_jacobi_from_zero<float>, gridn(ps->npts[0]), dim3(256), 0, st,
                          (float*)alt, (const float*)V.rho[0], (i64)ps->npts[0],
                          wpre ? wpre[0] : ps->w[0], op->st.coefs[0]);
        else
            launch_kernel(k_jacobi_from_zero<double>, gridn(ps->npts[0]), dim3(256), 0, st,
                          (double*)alt, (const double*)V.rho[0], (i64)ps->npts[0],
                          wpre ? wpre[0] : ps->w[0], op->st.coefs[0]);
        e = launch_error();
        std::swap(cur, alt);
        pre--;
        if (wpre) wpre++;
    } else if (l == 0 && zero0) {
        e = launch_fill(cur, V.dt, ps->npts[0], st);
    }
    if (e != hipSuccess) return e;
    if (l < m->L) {
        // below level 0, phi_l was zeroed by the restriction onto it
        e = relax_v(ps, V, l, pre, cur, alt, st, nullptr, l > 0, wpre);
        if (e != hipSuccess) return e;
        if (ps->fuse_restrict && resid_restrict_ok(op->st, op->g, ps->rst[l]->t)) {
            e = launch_resid_restrict(op->st, V.dt, cur, V.rho[l], V.rho[l + 1], V.phiA[l + 1],
                                      op->g, ps->rst[l]->t, st);
        } else {
            e = launch_stencil(op->st, V.dt, ST_RESID, cur, V.res[l], V.rho[l], op->g, ph, 1, 0.0,
                               nullptr, nullptr, true, st);
            if (e == hipSuccess)
                e = launch_transfer(ps->rst[l]->t, V.dt, TR_STORE, V.res[l], V.rho[l + 1],
                                    V.phiA[l + 1], ph, 1, st);
        }
        if (e != hipSuccess) return e;
        void* coarse = nullptr;
        e = emit_v(ps, m, V, l + 1, false, coarse, st);
        if (e != hipSuccess) return e;
        if (l > 0 && ps->fold_correction && ps->fuse_jacobi && jacobi_fused_ok(op->st, op->g) &&
            m->post[l] >= jacobi_fused_sweeps(op->g) && ps->itp[l]->t.K == 2) {
            e = relax_v(ps, V, l, m->post[l], cur, alt, st, coarse, false, wpost);
            res = cur;
            return e;
        }
        if (l == 0 && ps->fold_correction && m->post[0] >= 1 &&
            jacobi_corr_ok(op->st, op->g, ps->itp[0]->t)) {
            // phi_0 -= I phi_1 inside the first post-smoothing sweep
            e = relax_v(ps, V, l, m->post[l], cur, alt, st, coarse, false, wpost);
            res = cur;
            return e;
        }
        e = launch_transfer(ps->itp[l]->t, V.dt, TR_SUBTRACT, coarse, cur, nullptr, ph, 1, st);
        if (e != hipSuccess) return e;
    }
    e = relax_v(ps, V, l, m->post[l], cur, alt, st, nullptr, false, wpost);
    res = cur;
    return e;
}

// the whole inner V-cycle from level 0 (its result lands in phiA[0]: level
// 0 does an even number of single sweeps)
hipError_t emit_top(gpaw_poisson* ps, MixedState* m, int pr, bool zero0, hipStream_t st)
{
    void* res = nullptr;
    return emit_v(ps, m, m->v[pr], 0, zero0, res, st);
}

void* z_of(MixedState* m, int pr) { return m->v[pr].phiA[0]; }   // even level-0 sweep count

template <class F>
int capture_graph(hipGraphExec_t* out, F&& body)
{
    return gpaw_check(build_graph(out, body), "mixed: graph");
}

int alloc_bytes(void** p, size_t n, std::vector<void*>& owned, const char* what)
{
    int rc = gpaw_check(hipMalloc(p, n > 0 ? n : 1), what);
    if (!rc) owned.push_back(*p);
    return rc;
}

// The FP32 hierarchy: all levels in FP32; the fused coarse levels keep FP64
// arrays in LDS with an FP32 boundary.
int setup_v32(gpaw_poisson* ps, MixedState* m)
{
    VLevels& V = m->v[0];
    V.dt = DT_F32;
    int rc = GPAW_OK;
    for (int l = 0; l <= ps->L && !rc; l++) {
        void *a = nullptr, *b = nullptr, *r = nullptr, *x = nullptr;
        const size_t n = sizeof(float) * ps->npts[l];
        rc = alloc_bytes(&a, n, V.owned, "mixed: level arrays");
        if (!rc) rc = alloc_bytes(&b, n, V.owned, "mixed: level arrays");
        if (!rc) rc = alloc_bytes(&r, n, V.owned, "mixed: level arrays");
        if (!rc) rc = alloc_bytes(&x, n, V.owned, "mixed: level arrays");
        V.phiA.push_back(a);
        V.phiB.push_back(b);
        V.rho.push_back(r);
        V.res.push_back(x);
    }
    // the fused FP64 coarse solve needs its arrays in LDS (FP32 boundary)
    if (!rc && ps->lc > 0 && ps->h_coarse.lds) {
        CoarsePlan P = ps->h_coarse;
        P.lev[0].rho = (double*)V.rho[ps->lc];
        P.lev[0].phiA = (double*)V.phiA[ps->lc];
        P.lev[0].phiB = (double*)V.phiB[ps->lc];
        V.h_coarse = P;
        rc = gpaw_check(hipMalloc((void**)&V.d_coarse, sizeof P), "mixed: coarse plan");
        if (!rc) {
            V.own_coarse = true;
            rc = gpaw_check(hipMemcpy(V.d_coarse, &P, sizeof P, hipMemcpyHostToDevice),
                            "mixed: coarse plan");
        }
        if (!rc) V.lc = ps->lc;
    }
    return rc;
}

// The FP64 hierarchy: level 0 reads the FP64 residual itself and sweeps
// between two arrays of its own; the levels below reuse the plan's arrays
// and its fused FP64 coarse kernel.
int setup_v64(gpaw_poisson* ps, MixedState* m)
{
    VLevels& V = m->v[1];
    V.dt = DT_F64;
    const size_t n0 = sizeof(double) * ps->npts[0];
    void *a = nullptr, *b = nullptr, *x = nullptr;
    int rc = alloc_bytes(&a, n0, V.owned, "mixed: FP64 level-0 arrays");
    if (!rc) rc = alloc_bytes(&b, n0, V.owned, "mixed: FP64 level-0 arrays");
    if (!rc) rc = alloc_bytes(&x, n0, V.owned, "mixed: FP64 level-0 arrays");
    if (rc) return rc;
    V.phiA.push_back(a);
    V.phiB.push_back(b);
    V.rho.push_back(ps->res[0]);
    V.res.push_back(x);
    for (int l = 1; l <= ps->L; l++) {
        V.phiA.push_back(ps->phiA[l]);
        V.phiB.push_back(ps->phiB[l]);
        V.rho.push_back(ps->rho[l]);
        V.res.push_back(ps->res[l]);
    }
    if (ps->lc > 0) {
        V.d_coarse = ps->d_coarse;
        V.h_coarse = ps->h_coarse;
        V.lc = ps->lc;
    }
    return GPAW_OK;
}

std::vector<double> chebyshev_weights(const Stencil& s, double alpha, int k);
void mixed_level_schedule(const gpaw_poisson* ps, int l, int& pre, int& post,
                          std::vector<double>& wpre, std::vector<double>& wpost);

// The inner V-cycle's level-0 smoothing: 2 pre + 2 post Jacobi sweeps whose
// weights are the Chebyshev smoother for the top 1/7 of the spectrum of
// D^-1 A, post-smoothing in the reverse order, so that the V-cycle is a
// symmetric preconditioner (the sweeps' factors commute).  Measured on the
// 159^3 mixed-precision solve: 11 CG steps, 3.04 ms, against 12-14 steps,
// 3.6-3.9 ms for 2+2 or 3+3 sweeps of the plan's weight 2/3 (intervals
// lmax/3 .. lmax/12 and 1+1, 3+3 sweeps measured; lmax/6 .. lmax/10 with
// 2+2 were best).  GPAW_AMD_MIXED_SMOOTH=pre,post and GPAW_AMD_MIXED_CHEB=
// alpha (0: the plan's weight for every sweep) override it.
void mixed_schedule(const gpaw_poisson* ps, int& pre, int& post, std::vector<double>& wpre,
                    std::vector<double>& wpost)
{
    int s0 = 2, s1 = 2;
    double alpha = 7.0;
    if (const char* env = getenv("GPAW_AMD_MIXED_SMOOTH")) sscanf(env, "%d,%d", &s0, &s1);
    if (const char* env = getenv("GPAW_AMD_MIXED_CHEB")) alpha = atof(env);
    s0 = std::max(0, std::min(s0, MIXED_MAX_SWEEPS));
    s1 = std::max(0, std::min(s1, MIXED_MAX_SWEEPS - 1));
    if ((s0 + s1) % 2) s1++;             // the result must land in phiA[0]
    pre = s0;
    post = s1;
    if (ps->L == 0) post = s0 + s1;
    wpre.clear();
    wpost.clear();
    if (alpha > 1.0 && ps->L > 0 && ps->ops[0]->st.star_r > 0) {
        wpre = chebyshev_weights(ps->ops[0]->st, alpha, pre);
        wpost = chebyshev_weights(ps->ops[0]->st, alpha, post);
        std::reverse(wpost.begin(), wpost.end());
    }
}

// Levels below 0: the plan's sweeps and weight, except level 1 when its
// fused launches do two sweeps (a level whose 12^3 boxes fill the chip,
// e.g. 79^3 under 159^3): there 2 + 2 Chebyshev-weighted sweeps (the top
// 1/6 of the spectrum) save two of its four launches per V-cycle at the
// same CG step count (159^3: 3.09 -> 2.80 ms; on smaller grids, whose level
// 1 takes four sweeps per launch anyway, 4 + 4 stay: 2 + 2 added a step on
// 96^3; 2 + 2 on levels 2 or 3 of 159^3 took 12 steps).
// GPAW_AMD_MIXED_L<l>=pre,post,alpha overrides level l below the fused
// coarse levels (alpha <= 1: the plan's weight).
void mixed_level_schedule(const gpaw_poisson* ps, int l, int& pre, int& post,
                          std::vector<double>& wpre, std::vector<double>& wpost)
{
    pre = l < ps->L ? ps->pre[l] : 0;
    post = ps->post[l];
    wpre.clear();
    wpost.clear();
    if (l < 1 || l >= ps->L || (ps->lc >= 0 && l >= ps->lc) || ps->ops[l]->st.star_r <= 0) return;
    int a = 0, b = 0;
    double alpha = 0.0;
    if (l == 1 && jacobi_fused_ok(ps->ops[1]->st, ps->ops[1]->g) &&
        jacobi_fused_sweeps(ps->ops[1]->g) == 2) {
        a = b = 2;
        alpha = 6.0;
    }
    char name[32];
    snprintf(name, sizeof name, "GPAW_AMD_MIXED_L%d", l);
    if (const char* env = getenv(name)) sscanf(env, "%d,%d,%lf", &a, &b, &alpha);
    if (a <= 0 || b <= 0 || (a + b) % 2 || a > MIXED_MAX_SWEEPS || b > MIXED_MAX_SWEEPS) return;
    pre = a;
    post = b;
    if (alpha > 1.0) {
        wpre = chebyshev_weights(ps->ops[l]->st, alpha, a);
        wpost = chebyshev_weights(ps->ops[l]->st, alpha, b);
        std::reverse(wpost.begin(), wpost.end());
    }
}

int setup(gpaw_poisson* ps, int pr)
{
    if (!ps->mixed) {
        MixedState* m = new MixedState();
        ps->mixed = m;
        m->L = ps->L;
        m->pre = ps->pre;
        m->post = ps->post;
        mixed_schedule(ps, m->pre[0], m->post[0], m->w0pre, m->w0post);
        m->wpre.assign(ps->L + 1, {});
        m->wpost.assign(ps->L + 1, {});
        for (int l = 1; l <= ps->L; l++)
            mixed_level_schedule(ps, l, m->pre[l], m->post[l], m->wpre[l], m->wpost[l]);
        const size_t n0 = sizeof(double) * ps->npts[0];
        int rc = gpaw_check(hipMalloc((void**)&m->p, n0), "mixed: p");
        if (!rc) rc = gpaw_check(hipMalloc((void**)&m->q, n0), "mixed: q");
        if (!rc) rc = gpaw_check(hipMalloc((void**)&m->S, sizeof(double) * S_N), "mixed: scalars");
        if (!rc)
            rc = gpaw_check(hipHostMalloc((void**)&m->h_S, sizeof(double) * S_N), "mixed: pinned");
        if (!rc)
            rc = gpaw_check(hipMalloc((void**)&m->partial, sizeof(double) * 4 * NB),
                            "mixed: partials");
        if (!rc) rc = gpaw_check(hipMemset(m->S, 0, sizeof(double) * S_N), "mixed: scalars");
        if (rc) return rc;
    }
    MixedState* m = ps->mixed;
    if (m->ready[pr]) return GPAW_OK;
    int rc = pr ? setup_v64(ps, m) : setup_v32(ps, m);
    if (!rc) m->ready[pr] = true;
    return rc;
}

void destroy_graphs(MixedState* m, int pr)
{
    for (hipGraphExec_t* g : {&m->g_vcycle[pr], &m->g_vwarm[pr], &m->g_a[pr], &m->g_b[pr]}) {
        if (*g) hipGraphExecDestroy(*g);
        *g = nullptr;
    }
    for (int t = 0; t < 2; t++)
        for (hipGraphExec_t& g : m->g_ba[pr][t]) {
            if (g) hipGraphExecDestroy(g);
            g = nullptr;
        }
}

// PCG step part a: q = L p, alpha, phi += alpha p, d -= alpha q, |d|^2.
// |d|^2 (S[DD]) is only read by the host after a launch, so only the last
// step of a launch sums its partials (`last`).
hipError_t emit_a(gpaw_poisson* ps, MixedState* m, int pr, hipStream_t st, bool last = true)
{
    const i64 n = ps->npts[0];
    gpaw_operator* op = ps->ops[0];
    Phases ph;
    gpaw_phases(nullptr, ph);
    hipError_t e;
    const double* pqp = nullptr;
    int npq = 0;
    if (stencil_dot_ok(op->st, DT_F64, op->g, 1)) {   // q = L p with the partials of p.q
        int np = 0;
        e = launch_stencil(op->st, DT_F64, ST_APPLY, m->p, m->q, nullptr, op->g, ph, 1, 0.0,
                           op->partial, &np, true, st);
        if (e != hipSuccess) return e;
        pqp = op->partial;   // summed inside the update
        npq = np;
    } else {
        e = launch_stencil(op->st, DT_F64, ST_APPLY, m->p, m->q, nullptr, op->g, ph, 1, 0.0,
                           nullptr, nullptr, true, st);
        if (e != hipSuccess) return e;
        launch_kernel(k_dot, gridn(n), dim3(256), 0, st, m->p, m->q, n, m->partial + PART_PQ);
        launch_kernel(k_red1, dim3(1), dim3(256), 0, st, m->partial + PART_PQ,
                      (int)gridn(n).x, m->S, (int)S_PQ);
    }
    // dz: the d.z partials of the step's k_dot_dz_zq
    if (pr == 0)
        launch_kernel(k_pcg_update<float>, gridn(n), dim3(256), 0, st, ps->phiA[0], ps->res[0],
                      m->p, m->q, (float*)m->v[0].rho[0], m->S, n, m->partial + PART_DD, pqp,
                      npq, (const double*)m->partial, (int)gridn(n).x);
    else
        launch_kernel(k_pcg_update<double>, gridn(n), dim3(256), 0, st, ps->phiA[0], ps->res[0],
                      m->p, m->q, (double*)nullptr, m->S, n, m->partial + PART_DD, pqp, npq,
                      (const double*)m->partial, (int)gridn(n).x);
    if (last)
        launch_kernel(k_red1, dim3(1), dim3(256), 0, st, m->partial + PART_DD, (int)gridn(n).x,
                      m->S, (int)S_DD);
    return launch_error();
}

// z, beta (flexible), p = z + beta p after the inner V-cycle; first: p = z
template <class T>
void emit_dz_p(gpaw_poisson* ps, MixedState* m, int pr, int first, hipStream_t st,
               double* hist = nullptr)
{
    const i64 n = ps->npts[0];
    launch_kernel(k_dot_dz_zq<T>, gridn(n), dim3(256), 0, st, ps->res[0],
                  (const T*)z_of(m, pr), m->q, n, m->partial);
    launch_kernel(k_p_update<T>, gridn(n), dim3(256), 0, st, m->p, (const T*)z_of(m, pr),
                  m->S, n, first, (const double*)m->partial + gridn(n).x,
                  (int)gridn(n).x, (const double*)m->partial + PART_DD, (int)gridn(n).x, hist);
}

// PCG step part b: z = V(d), beta (flexible), p = z + beta p; `hist`: where
// the previous step's |d|^2 goes (nullptr: not summed)
hipError_t emit_b(gpaw_poisson* ps, MixedState* m, int pr, hipStream_t st, double* hist = nullptr)
{
    hipError_t e = emit_top(ps, m, pr, true, st);
    if (e != hipSuccess) return e;
    if (pr == 0)
        emit_dz_p<float>(ps, m, pr, 0, st, hist);
    else
        emit_dz_p<double>(ps, m, pr, 0, st, hist);
    return launch_error();
}

int capture_all(gpaw_poisson* ps, int pr)
{
    MixedState* m = ps->mixed;
    if (m->g_b[pr]) return GPAW_OK;   // the last one built: all four exist
    int rc = capture_graph(&m->g_vcycle[pr],
                           [&](hipStream_t st) { return emit_top(ps, m, pr, true, st); });
    if (!rc)
        rc = capture_graph(&m->g_vwarm[pr],
                           [&](hipStream_t st) { return emit_top(ps, m, pr, false, st); });
    if (!rc) rc = capture_graph(&m->g_a[pr], [&](hipStream_t st) { return emit_a(ps, m, pr, st); });
    if (!rc) rc = capture_graph(&m->g_b[pr], [&](hipStream_t st) { return emit_b(ps, m, pr, st); });
    if (!rc)
        for (int t = 0; t < 2; t++) m->g_ba[pr][t].assign(MIXED_CHUNK_MAX, nullptr);
    if (rc) destroy_graphs(m, pr);   // retried from scratch on the next call
    return rc;
}

// |L phi - rho|^2 (FP64) into S[slot]; phi, rho and the iteration's
// residual d are left as they are
hipError_t emit_true_norm(gpaw_poisson* ps, MixedState* m, int slot, hipStream_t st)
{
    gpaw_operator* op = ps->ops[0];
    Phases ph;
    gpaw_phases(nullptr, ph);
    int np = 0;
    hipError_t e = launch_stencil(op->st, DT_F64, ST_RESID, ps->phiA[0], nullptr, ps->rho[0],
                                  op->g, ph, 1, 0.0, op->partial, &np, false, st);
    if (e == hipSuccess) e = launch_reduce_partials(op->partial, np, 1.0, m->S + slot, st);
    return e;
}

// k PCG steps (b, a) in one graph; `fin`: then the true residual's norm
// into S[DT] (a chunk expected to reach the tolerance)
int run_ba(gpaw_poisson* ps, MixedState* m, int pr, int k, bool fin, hipStream_t st)
{
    hipGraphExec_t& g = m->g_ba[pr][fin ? 1 : 0][k - 1];
    if (!g) {
        int rc = capture_graph(&g, [&](hipStream_t s) -> hipError_t {
            hipError_t e = hipSuccess;
            for (int i = 0; i < k && e == hipSuccess; i++) {
                e = emit_b(ps, m, pr, s, i > 0 ? m->S + S_HIST + i - 1 : nullptr);
                if (e == hipSuccess) e = emit_a(ps, m, pr, s, i == k - 1);
            }
            if (e == hipSuccess && fin) e = emit_true_norm(ps, m, S_DT, s);
            return e;
        });
        if (rc) return rc;
    }
    return gpaw_check(hipGraphLaunch(g, st), "mixed: iterations");
}

// Weights of k Jacobi sweeps whose product polynomial prod_i (1 - w_i D^-1 A)
// is the degree-k Chebyshev polynomial that is smallest on [lmax / alpha,
// lmax] (the oscillatory modes the coarse grid does not correct), lmax the
// largest eigenvalue of D^-1 A for the star stencil s (its symbol's maximum,
// a sum of one-axis terms).
std::vector<double> chebyshev_weights(const Stencil& s, double alpha, int k)
{
    const double c0 = s.coefs[0];
    double lmax = 1.0;
    for (int a = 0; a < 3; a++) {
        double best = -1e300;
        for (int t = 0; t <= 512; t++) {
            const double th = M_PI * t / 512.0;
            double g = 0.0;
            for (int c = 1; c < s.nc; c++) {
                const int* o = s.off[c];
                const int other = (a + 1) % 3, third = (a + 2) % 3;
                if (o[other] == 0 && o[third] == 0 && o[a] != 0) g += s.coefs[c] * std::cos(o[a] * th);
            }
            best = std::max(best, g / c0);
        }
        lmax += best;
    }
    const double b = lmax, lo = lmax / alpha;
    std::vector<double> w(k);
    for (int i = 0; i < k; i++) {
        const double root = 0.5 * (b + lo) + 0.5 * (b - lo) * std::cos(M_PI * (2 * i + 1) / (2.0 * k));
        w[i] = 1.0 / root;
    }
    if (getenv("GPAW_AMD_MIXED_LOG"))
        for (int i = 0; i < k; i++) fprintf(stderr, "mixed: lmax %.4f weight %d = %.4f\n", lmax, i, w[i]);
    return w;
}

// GPAW_AMD_MIXED_LOG=1: the chunk schedule of every solve on stderr
bool mixed_log()
{
    static const bool on = [] {
        const char* e = getenv("GPAW_AMD_MIXED_LOG");
        return e && atoi(e) > 0;
    }();
    return on;
}

// S[DZ] = 0: the next direction update restarts the recursion (p = z)
int mark_restart(MixedState* m, hipStream_t st)
{
    return gpaw_check(hipMemsetAsync(m->S + S_DZ, 0, sizeof(double), st), "mixed: restart");
}

int read_scalar(MixedState* m, int slot, hipStream_t st, double* v)
{
    hipError_t e = hipMemcpyAsync(m->h_S, m->S, sizeof(double) * S_N, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return gpaw_check(e, "mixed: readback");
    *v = m->h_S[slot];
    return GPAW_OK;
}

// d = L phi - rho (FP64, stored in res[0]) with |d|^2 into S[slot]; the
// FP32 inner V-cycle also gets d32 = d
int residual64(gpaw_poisson* ps, MixedState* m, int pr, hipStream_t st, int slot = S_DD)
{
    gpaw_operator* op = ps->ops[0];
    Phases ph;
    gpaw_phases(nullptr, ph);
    int np = 0;
    hipError_t e = launch_stencil(op->st, DT_F64, ST_RESID, ps->phiA[0], ps->res[0], ps->rho[0],
                                  op->g, ph, 1, 0.0, op->partial, &np, true, st);
    if (e == hipSuccess) e = launch_reduce_partials(op->partial, np, 1.0, m->S + slot, st);
    if (e == hipSuccess && pr == 0)
        e = launch_cast(m->v[0].rho[0], DT_F32, ps->res[0], DT_F64, ps->npts[0], st);
    return gpaw_check(e, "mixed: residual");
}

}  // namespace

void gpaw_mixed_free(MixedState* m)
{
    if (!m) return;
    for (int pr = 0; pr < 2; pr++) {
        destroy_graphs(m, pr);
        for (void* p : m->v[pr].owned) hipFree(p);
        if (m->v[pr].own_coarse && m->v[pr].d_coarse) hipFree(m->v[pr].d_coarse);
    }
    for (double* p : {m->p, m->q, m->S, m->partial})
        if (p) hipFree(p);
    if (m->h_S) hipHostFree(m->h_S);
    delete m;
}

extern "C" int gpaw_poisson_precondition(gpaw_poisson* ps, const double* r, double* z,
                                         int inner_precision, void* stream)
{
    if (!ps || !r || !z || inner_precision < 0 || inner_precision > 1)
        return gpaw_fail(GPAW_EINVAL, "precondition: bad argument");
    if (ps->method != 2) return gpaw_fail(GPAW_EUNSUPPORTED, "precondition: Jacobi smoothing only");
    if (r == z) return gpaw_fail(GPAW_EINVAL, "precondition: r and z must differ");
    const int pr = inner_precision;
    hipStream_t st = (hipStream_t)stream;
    int rc = setup(ps, pr);
    if (rc) return rc;
    MixedState* m = ps->mixed;
    const i64 n = ps->npts[0];
    // the right-hand side where the inner V-cycle reads it: the FP32
    // hierarchy's level-0 rho, or (FP64) the plan's level-0 residual array
    hipError_t e = pr == 0 ? launch_cast(m->v[0].rho[0], DT_F32, r, DT_F64, n, st)
                           : launch_copy(ps->res[0], r, DT_F64, n, st);
    if (e != hipSuccess) return gpaw_check(e, "precondition: load");
    if (!m->tuned[pr] && autotune_set(-1)) {
        m->tuned[pr] = 1;
        e = emit_top(ps, m, pr, true, st);
        if (e != hipSuccess) return gpaw_check(e, "precondition: tune");
    }
    rc = capture_all(ps, pr);
    if (rc) return rc;
    rc = gpaw_check(hipGraphLaunch(m->g_vcycle[pr], st), "precondition: V-cycle");
    if (rc) return rc;
    e = pr == 0 ? launch_cast(z, DT_F64, z_of(m, pr), DT_F32, n, st)
                : launch_copy(z, z_of(m, pr), DT_F64, n, st);
    return gpaw_check(e, "precondition: store");
}

extern "C" int gpaw_poisson_solve_mixed(gpaw_poisson* ps, double* phi, const double* rho,
                                        double rel_tol, int maxiter, int mode, int inner,
                                        int* niter, double* relres, void* stream)
{
    // rho == nullptr: the density gpaw_poisson_neutralize left in the plan
    if (!ps || !phi || (!rho && !ps->rho_loaded) || maxiter < 1 || rel_tol <= 0.0 || mode < 0 || mode > 7 ||
        inner < 1)
        return gpaw_fail(GPAW_EINVAL, "mixed: bad argument");
    if (ps->method != 2) return gpaw_fail(GPAW_EUNSUPPORTED, "mixed: Jacobi smoothing only");
    const int pcg = mode & 1;   // else iterative refinement
    const int pr = (mode >> 1) & 1;   // inner V-cycle precision: 0 FP32, 1 FP64
    const bool history = (mode & 4) != 0;   // first chunk from the last solve
    hipStream_t st = (hipStream_t)stream;
    int rc = setup(ps, pr);
    if (rc) return rc;
    MixedState* m = ps->mixed;
    if (!m->tuned[pr] && autotune_set(-1)) {
        // one eager pass of the graph bodies on the plan's own buffers (all
        // reloaded below), so that the graphs record tuned launch shapes
        m->tuned[pr] = 1;
        hipError_t e = emit_top(ps, m, pr, true, st);
        if (e == hipSuccess) e = emit_a(ps, m, pr, st);
        rc = gpaw_check(e, "mixed: tune");
    }
    if (!rc) rc = capture_all(ps, pr);
    if (rc) return rc;
    const i64 n = ps->npts[0];
    hipError_t e = rho ? launch_copy(ps->rho[0], rho, DT_F64, n, st) : hipSuccess;
    ps->rho_loaded = false;
    if (e == hipSuccess) e = launch_copy(ps->phiA[0], phi, DT_F64, n, st);
    if (e == hipSuccess) {   // |rho|^2 for the relative criterion
        launch_kernel(k_dot, gridn(n), dim3(256), 0, st, ps->rho[0], ps->rho[0], n, m->partial);
        launch_kernel(k_red1, dim3(1), dim3(256), 0, st, m->partial, (int)gridn(n).x, m->S,
                      (int)S_RR);
        e = launch_error();
    }
    if (e != hipSuccess) return gpaw_check(e, "mixed: load");
    double rr = 0.0, dd = 0.0, target = 0.0;
    int it = 0;
    if (pcg) {
        // d = L phi - rho with |d|^2 into S[D0] and a restart marker, then
        // the first chunk with no readback in between (on a zero residual
        // its steps change nothing).  With `history` the first chunk is as
        // long as the previous solve on this plan took and ends with the
        // true residual's norm, so a repeated solve of a similar density
        // reads back once.  Later chunks follow the observed convergence
        // rate; a chunk may run a few steps past the tolerance, which only
        // makes the result more accurate.
        rc = residual64(ps, m, pr, st, S_D0);
        if (!rc) rc = mark_restart(m, st);
        if (rc) return rc;
        const int h = (history && m->hint_tol[pr] == rel_tol) ? m->hint[pr] : 0;
        int k = h > 0 ? std::min(h, MIXED_CHUNK_MAX) : 3;
        bool fin = h > 0 && h <= MIXED_CHUNK_MAX;
        // |d|^2 after every step since the last (re)start (index 0: the
        // start) and the step count of that start
        std::vector<double> hist;
        int base = 0, cross = 0;   // cross: steps after which |d| first met the target
        bool first = true;
        for (;;) {
            k = std::min(k, maxiter - it);
            if (k < 1) break;
            rc = run_ba(ps, m, pr, k, fin, st);
            if (!rc) rc = read_scalar(m, S_DD, st, &dd);
            if (rc) break;
            if (first) {   // |rho|^2 and the starting |d|^2 came back too
                rr = m->h_S[S_RR];
                target = rel_tol * rel_tol * rr;
                hist.push_back(m->h_S[S_D0]);
                first = false;
            }
            for (int j = 0; j < k; j++) {
                const double v = j < k - 1 ? m->h_S[S_HIST + j] : dd;
                hist.push_back(v);
                if (!cross && v <= target) cross = base + (int)hist.size() - 1;
            }
            it += k;
            if (mixed_log())
                fprintf(stderr, "mixed: chunk %d%s -> %d steps, |d|^2/target %.3g%s\n", k,
                        fin ? " +true" : "", it, dd / target,
                        fin ? "" : " (no true residual)");
            if (fin && mixed_log()) fprintf(stderr, "mixed:   true/target %.3g\n", m->h_S[S_DT] / target);
            if (!std::isfinite(dd)) break;
            // the true residual L phi - rho decides; the recursion's |d| can
            // drift from it
            bool have_true = fin;
            double dt = fin ? m->h_S[S_DT] : 0.0;
            if (!fin && dd <= target) {
                rc = gpaw_check(emit_true_norm(ps, m, S_DT, st), "mixed: residual");
                if (!rc) rc = read_scalar(m, S_DT, st, &dt);
                if (rc) break;
                have_true = true;
            }
            if (have_true && (!std::isfinite(dt) || dt <= target)) {
                dd = dt;
                break;
            }
            if (it >= maxiter) break;
            if (have_true && dd <= target) {
                // drifted: restart from the true residual (p = z)
                rc = residual64(ps, m, pr, st);
                if (!rc) rc = mark_restart(m, st);
                if (rc) break;
                dd = dt;
                hist.assign(1, dt);
                base = it;
                cross = 0;
                k = 3;
                fin = false;
                continue;
            }
            // the steps still needed at the rate of the last (up to) two
            // steps, at most 8 per chunk: the rate of the first steps
            // underestimates the later ones (16 steps predicted where 11
            // sufficed on 159^3), and a step run past the tolerance costs
            // more than a readback
            k = MIXED_PREDICT_MAX;
            fin = false;
            const int ns = (int)hist.size() - 1, w = std::min(2, ns);
            const double old = hist[ns - w];
            if (w > 0 && dd > target && dd < old && old > 0) {
                const double rate = std::pow(dd / old, 1.0 / w);
                if (rate > 0 && rate < 1) {
                    const double need = std::ceil(std::log(target / dd) / std::log(rate));
                    if (need >= 1 && need <= MIXED_PREDICT_MAX) {
                        k = (int)need;
                        fin = true;
                    }
                }
            }
        }
        if (!rc && dd <= target) {
            m->hint[pr] = cross > 0 ? cross : it;
            m->hint_tol[pr] = rel_tol;
        }
    } else {
        // |rho|^2 and the first |d|^2 come back in one readback
        rc = residual64(ps, m, pr, st);
        if (!rc) rc = read_scalar(m, S_DD, st, &dd);
        if (rc) return rc;
        rr = m->h_S[S_RR];
        target = rel_tol * rel_tol * rr;
    }
    while (!pcg && !rc && !(dd <= target) && std::isfinite(dd) && it < maxiter) {
        rc = gpaw_check(hipGraphLaunch(m->g_vcycle[pr], st), "mixed: V-cycle");
        for (int k = 1; k < inner && !rc; k++)
            rc = gpaw_check(hipGraphLaunch(m->g_vwarm[pr], st), "mixed: V-cycle");
        if (!rc) {
            if (pr == 0)
                launch_kernel(k_sub_corr<float>, gridn(n), dim3(256), 0, st, ps->phiA[0],
                              (const float*)z_of(m, pr), n);
            else
                launch_kernel(k_sub_corr<double>, gridn(n), dim3(256), 0, st, ps->phiA[0],
                              (const double*)z_of(m, pr), n);
            rc = gpaw_check(launch_error(), "mixed: correction");
        }
        if (!rc) rc = residual64(ps, m, pr, st);
        if (!rc) rc = read_scalar(m, S_DD, st, &dd);
        it++;
    }
    if (rc) return rc;
    if (niter) *niter = it;
    if (relres) *relres = rr > 0 ? sqrt(dd / rr) : 0.0;
    if (!std::isfinite(dd))
        return gpaw_fail(GPAW_EDEVICE, "mixed-precision Poisson solver broke down (residual "
                                       "norm %g after %d iterations); phi not updated",
                         dd, it);
    rc = gpaw_check(launch_copy(phi, ps->phiA[0], DT_F64, n, st), "mixed: store");
    if (rc) return rc;
    if (!(dd <= target))
        return gpaw_fail(GPAW_ENOCONV, "mixed-precision Poisson solver did not reach %g in %d "
                                       "iterations",
                         rel_tol, maxiter);
    return GPAW_OK;
}

extern "C" int gpaw_poisson_mixed_weights(const gpaw_poisson* ps, int level, double* pre,
                                          int* npre, double* post, int* npost)
{
    if (!ps || !npre || !npost || level < 0 || level > ps->L)
        return gpaw_fail(GPAW_EINVAL, "mixed weights: bad argument");
    int a = 0, b = 0;
    std::vector<double> wa, wb;
    if (level == 0)
        mixed_schedule(ps, a, b, wa, wb);
    else
        mixed_level_schedule(ps, level, a, b, wa, wb);
    *npre = level < ps->L ? a : 0;
    *npost = b;
    for (int i = 0; i < *npre && pre; i++) pre[i] = wa.empty() ? ps->w[level] : wa[i];
    for (int i = 0; i < b && post; i++) post[i] = wb.empty() ? ps->w[level] : wb[i];
    return GPAW_OK;
}
