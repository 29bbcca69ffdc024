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
//   mode 0  iterative refinement: d = L phi - rho; e = V32^k(d); phi -= e
//   mode 1  flexible preconditioned conjugate gradients with one FP32
//           V-cycle as preconditioner (Polak-Ribiere beta, which tolerates
//           the preconditioner's FP32 rounding), i.e. iterative refinement
//           with a Krylov recurrence on the corrections.
// The inner V-cycle uses the reference hierarchy (gpaw/poisson.py:64-113)
// with symmetric level-0 smoothing (2 pre + 2 post by default) so that it is
// a symmetric preconditioner.  Stopping rule: ||d||_2 <= tol * ||rho||_2.
//
// Scalars (alpha, beta, dot products) stay on the device; the host reads
// one 8-byte norm per outer iteration.  The FP32 V-cycle and the per-
// iteration kernel sequences are captured as hipGraphs.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "poisson_internal.h"

using namespace gpaw;

struct MixedState {
    int L = 0;
    int lc = -1;                        // fused FP64 coarse level (-1: none)
    std::vector<int> pre, post;
    std::vector<float*> phiA, phiB, rho, res;
    CoarsePlan* d_coarse32 = nullptr;
    double *p = nullptr, *q = nullptr;  // FP64 search direction, L p
    double* S = nullptr;                // device scalars
    double* h_S = nullptr;              // pinned readback
    double* partial = nullptr;
    int nblk = 0;
    hipGraphExec_t g_vcycle = nullptr;  // one FP32 V-cycle from zero
    hipGraphExec_t g_vwarm = nullptr;   // one FP32 V-cycle from the last result
    hipGraphExec_t g_a = nullptr;       // PCG: q = Lp, alpha, update, |d|^2
    hipGraphExec_t g_b = nullptr;       // PCG: z = V32(d), beta, p = z + beta p
    // k PCG steps (b then a) per launch, k = 1 .. MIXED_CHUNK_MAX
    std::vector<hipGraphExec_t> g_ba;
    int tuned = 0;                      // launch shapes tuned (tuner on)
};

#define MIXED_CHUNK_MAX 8

// device scalars
enum { S_DZ = 0, S_PQ = 1, S_DD = 2, S_RR = 4, S_ALPHA = 5, S_N = 8 };

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

// alpha = dz / pq ; phi -= alpha p ; d -= alpha q ; d32 = d ; partials of
// |d|^2 into `partial`.  dz (the ndz partials at dzp) and pq (the npq
// partials at pqp, or S[PQ] when npq = 0) are summed by every block itself;
// block 0 keeps dz in S[DZ] (the next step's old dz) and alpha in S[ALPHA].
__global__ __launch_bounds__(256) void k_pcg_update(double* __restrict__ phi, double* __restrict__ d,
                                                    const double* __restrict__ p,
                                                    const double* __restrict__ q,
                                                    float* __restrict__ d32, double* __restrict__ S,
                                                    i64 n, double* __restrict__ partial,
                                                    const double* __restrict__ pqp, int npq,
                                                    const double* __restrict__ dzp, int ndz)
{
    __shared__ double red[256];
    const double pq = npq > 0 ? sum_partials(pqp, npq, red) : S[S_PQ];
    const double dz = sum_partials(dzp, ndz, red);
    const double alpha = dz / pq;
    double s = 0.0;
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += (i64)gridDim.x * 256) {
        phi[i] = phi[i] - alpha * p[i];
        const double di = d[i] - alpha * q[i];
        d[i] = di;
        d32[i] = (float)di;
        s += di * di;
    }
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S[S_ALPHA] = alpha;
        S[S_DZ] = dz;
    }
}

// partials of d.z and z.q (z from the FP32 preconditioner)
__global__ __launch_bounds__(256) void k_dot_dz_zq(const double* __restrict__ d,
                                                   const float* __restrict__ z,
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
// (kept by the previous update); first: p = z
__global__ __launch_bounds__(256) void k_p_update(double* __restrict__ p, const float* __restrict__ z,
                                                  const double* __restrict__ S, i64 n, int first,
                                                  const double* __restrict__ zqp, int nzq)
{
    __shared__ double red[256];
    const double zq = first ? 0.0 : sum_partials(zqp, nzq, red);
    const double beta = first ? 0.0 : -S[S_ALPHA] * zq / S[S_DZ];
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += (i64)gridDim.x * 256)
        p[i] = first ? (double)z[i] : (double)z[i] + beta * p[i];
}

// The first Jacobi sweep from phi = 0: every stencil term is zero, so the
// sweep is element-wise: the tile kernel's epilogue with a zero sum and a
// zero centre value (same operations, same bits).
__global__ __launch_bounds__(256) void k_jacobi_from_zero(float* __restrict__ out,
                                                          const float* __restrict__ rho, i64 n,
                                                          double w, double c0)
{
    for (i64 i = (i64)blockIdx.x * 256 + threadIdx.x; i < n; i += (i64)gridDim.x * 256)
        out[i] = st_epilogue<ST_JACOBI>(0.0f, 0.0f, rho[i], (float)w, (float)c0);
}

// phi -= e (FP64 accumulation of an FP32 correction)
__global__ __launch_bounds__(256) void k_sub_f32(double* __restrict__ phi, const float* __restrict__ e,
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
hipError_t relax32(gpaw_poisson* ps, MixedState* m, int l, int k, float*& cur, float*& alt,
                   hipStream_t st, const float* corr = nullptr, bool zero_in = false)
{
    Phases ph;
    gpaw_phases(nullptr, ph);
    gpaw_operator* op = ps->ops[l];
    // below level 0, runs of JF_SWEEPS sweeps are one fused launch (callers
    // follow `cur`, not the sweep parity)
    if (l > 0 && ps->fuse_jacobi && jacobi_fused_ok(op->st, op->g)) {
        for (; k >= JF_SWEEPS; k -= JF_SWEEPS) {
            hipError_t e = launch_jacobi_fused(op->st, DT_F32, cur, alt, m->rho[l], op->g,
                                               ps->w[l], st, corr,
                                               corr ? &ps->itp[l]->t : nullptr, zero_in);
            corr = nullptr;
            zero_in = false;
            if (e != hipSuccess) return e;
            float* t = cur;
            cur = alt;
            alt = t;
        }
    }
    for (int it = 0; it < k; it++) {
        hipError_t e = launch_stencil(op->st, DT_F32, ST_JACOBI, cur, alt, m->rho[l], op->g, ph,
                                      1, ps->w[l], nullptr, nullptr, true, st);
        if (e != hipSuccess) return e;
        float* t = cur;
        cur = alt;
        alt = t;
    }
    return hipSuccess;
}

// one FP32 V-cycle on level l (level 0: rho = d32, result in phiA[0]);
// `res` returns the buffer that holds level l's result
hipError_t emit32(gpaw_poisson* ps, MixedState* m, int l, bool zero0, float*& res,
                  hipStream_t st)
{
    Phases ph;
    gpaw_phases(nullptr, ph);
    if (l == m->lc) {   // the fused coarse kernel leaves it where single sweeps would
        const int kc = (l < m->L ? m->pre[l] : 0) + m->post[l];
        res = (kc % 2) ? m->phiB[l] : m->phiA[l];
        return launch_coarse_vcycle_f32(m->d_coarse32, ps->coarse_lds, st);
    }
    gpaw_operator* op = ps->ops[l];
    float* cur = m->phiA[l];
    float* alt = m->phiB[l];
    hipError_t e = hipSuccess;
    int pre = l < m->L ? m->pre[l] : 0;
    if (l == 0 && zero0 && pre > 0 && ps->fuse_jacobi) {
        // the first sweep from zero, element-wise, into alt (as a sweep would)
        launch_kernel(k_jacobi_from_zero, gridn(ps->npts[0]), dim3(256), 0, st, alt, m->rho[0],
                      (i64)ps->npts[0], ps->w[0], op->st.coefs[0]);
        e = launch_error();
        std::swap(cur, alt);
        pre--;
    } else if (l == 0 && zero0) {
        e = launch_fill(cur, DT_F32, ps->npts[0], st);
    }
    if (e != hipSuccess) return e;
    if (l < m->L) {
        // below level 0, phi_l was zeroed by the restriction onto it
        e = relax32(ps, m, l, pre, cur, alt, st, nullptr, l > 0);
        if (e != hipSuccess) return e;
        e = launch_stencil(op->st, DT_F32, ST_RESID, cur, m->res[l], m->rho[l], op->g, ph, 1, 0.0,
                           nullptr, nullptr, true, st);
        if (e != hipSuccess) return e;
        e = launch_transfer(ps->rst[l]->t, DT_F32, TR_STORE, m->res[l], m->rho[l + 1],
                            m->phiA[l + 1], ph, 1, st);
        if (e != hipSuccess) return e;
        float* coarse = nullptr;
        e = emit32(ps, m, l + 1, false, coarse, st);
        if (e != hipSuccess) return e;
        if (l > 0 && ps->fold_correction && ps->fuse_jacobi && m->post[l] >= JF_SWEEPS &&
            jacobi_fused_ok(op->st, op->g) && ps->itp[l]->t.K == 2) {
            e = relax32(ps, m, l, m->post[l], cur, alt, st, coarse);
            res = cur;
            return e;
        }
        e = launch_transfer(ps->itp[l]->t, DT_F32, TR_SUBTRACT, coarse, cur, nullptr, ph, 1, st);
        if (e != hipSuccess) return e;
    }
    e = relax32(ps, m, l, m->post[l], cur, alt, st);
    res = cur;
    return e;
}

// the whole FP32 V-cycle from level 0 (its result lands in phiA[0]: level 0
// does an even number of single sweeps)
hipError_t emit32_top(gpaw_poisson* ps, MixedState* m, bool zero0, hipStream_t st)
{
    float* res = nullptr;
    return emit32(ps, m, 0, zero0, res, st);
}

float* z_of(MixedState* m) { return m->phiA[0]; }   // even level-0 sweep count

template <class F>
int capture_graph(hipGraphExec_t* out, F&& body)
{
    return gpaw_check(build_graph(out, body), "mixed: graph");
}

int setup(gpaw_poisson* ps)
{
    if (ps->mixed) return GPAW_OK;
    MixedState* m = new MixedState();
    ps->mixed = m;
    m->L = ps->L;
    m->pre = ps->pre;
    m->post = ps->post;
    int s0 = 2, s1 = 2;
    if (const char* env = getenv("GPAW_AMD_MIXED_SMOOTH")) sscanf(env, "%d,%d", &s0, &s1);
    if ((s0 + s1) % 2) s1++;             // result must land in phiA[0]
    m->pre[0] = s0;
    m->post[0] = s1;
    if (ps->L == 0) m->post[0] = s0 + s1;
    int rc = GPAW_OK;
    for (int l = 0; l <= ps->L && !rc; l++) {
        float *a = nullptr, *b = nullptr, *r = nullptr, *x = nullptr;
        const size_t n = sizeof(float) * (ps->npts[l] > 0 ? ps->npts[l] : 1);
        rc = gpaw_check(hipMalloc((void**)&a, n), "mixed: level arrays");
        if (!rc) rc = gpaw_check(hipMalloc((void**)&b, n), "mixed: level arrays");
        if (!rc) rc = gpaw_check(hipMalloc((void**)&r, n), "mixed: level arrays");
        if (!rc) rc = gpaw_check(hipMalloc((void**)&x, n), "mixed: level arrays");
        m->phiA.push_back(a);
        m->phiB.push_back(b);
        m->rho.push_back(r);
        m->res.push_back(x);
    }
    const size_t n0 = sizeof(double) * ps->npts[0];
    if (!rc) rc = gpaw_check(hipMalloc((void**)&m->p, n0), "mixed: p");
    if (!rc) rc = gpaw_check(hipMalloc((void**)&m->q, n0), "mixed: q");
    if (!rc) rc = gpaw_check(hipMalloc((void**)&m->S, sizeof(double) * S_N), "mixed: scalars");
    if (!rc) rc = gpaw_check(hipHostMalloc((void**)&m->h_S, sizeof(double) * S_N), "mixed: pinned");
    if (!rc) rc = gpaw_check(hipMalloc((void**)&m->partial, sizeof(double) * 4 * NB), "mixed: partials");
    if (!rc) rc = gpaw_check(hipMemset(m->S, 0, sizeof(double) * S_N), "mixed: scalars");
    // the fused FP64 coarse solve needs its arrays in LDS (FP32 boundary)
    if (!rc && ps->lc > 0 && ps->h_coarse.lds) {
        CoarsePlan P = ps->h_coarse;
        P.lev[0].rho = (double*)m->rho[ps->lc];
        P.lev[0].phiA = (double*)m->phiA[ps->lc];
        P.lev[0].phiB = (double*)m->phiB[ps->lc];
        rc = gpaw_check(hipMalloc((void**)&m->d_coarse32, sizeof P), "mixed: coarse plan");
        if (!rc) rc = gpaw_check(hipMemcpy(m->d_coarse32, &P, sizeof P, hipMemcpyHostToDevice),
                                 "mixed: coarse plan");
        if (!rc) m->lc = ps->lc;
    }
    return rc;
}

void destroy_graphs(MixedState* m)
{
    for (hipGraphExec_t* g : {&m->g_vcycle, &m->g_vwarm, &m->g_a, &m->g_b}) {
        if (*g) hipGraphExecDestroy(*g);
        *g = nullptr;
    }
    for (hipGraphExec_t& g : m->g_ba) {
        if (g) hipGraphExecDestroy(g);
        g = nullptr;
    }
}

// PCG step part a: q = L p, alpha, phi += alpha p, d -= alpha q, |d|^2.
// |d|^2 (S[DD]) is only read by the host after a launch, so only the last
// step of a launch sums its partials (`last`).
hipError_t emit_a(gpaw_poisson* ps, MixedState* m, hipStream_t st, bool last = true)
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
    launch_kernel(k_pcg_update, gridn(n), dim3(256), 0, st, ps->phiA[0], ps->res[0], m->p, m->q,
                  m->rho[0], m->S, n, m->partial + PART_DD, pqp, npq, (const double*)m->partial,
                  (int)gridn(n).x);
    if (last)
        launch_kernel(k_red1, dim3(1), dim3(256), 0, st, m->partial + PART_DD, (int)gridn(n).x,
                      m->S, (int)S_DD);
    return launch_error();
}

// PCG step part b: z = V32(d), beta (flexible), p = z + beta p
hipError_t emit_b(gpaw_poisson* ps, MixedState* m, hipStream_t st)
{
    const i64 n = ps->npts[0];
    hipError_t e = emit32_top(ps, m, true, st);
    if (e != hipSuccess) return e;
    launch_kernel(k_dot_dz_zq, gridn(n), dim3(256), 0, st, ps->res[0], z_of(m), m->q, n,
                  m->partial);
    launch_kernel(k_p_update, gridn(n), dim3(256), 0, st, m->p, z_of(m), (const double*)m->S, n,
                  0, (const double*)m->partial + gridn(n).x, (int)gridn(n).x);
    return launch_error();
}

int capture_all(gpaw_poisson* ps)
{
    MixedState* m = ps->mixed;
    if (m->g_b) return GPAW_OK;   // the last one built: all four exist
    int rc = capture_graph(&m->g_vcycle,
                           [&](hipStream_t st) { return emit32_top(ps, m, true, st); });
    if (!rc)
        rc = capture_graph(&m->g_vwarm,
                           [&](hipStream_t st) { return emit32_top(ps, m, false, st); });
    if (!rc) rc = capture_graph(&m->g_a, [&](hipStream_t st) { return emit_a(ps, m, st); });
    if (!rc) rc = capture_graph(&m->g_b, [&](hipStream_t st) { return emit_b(ps, m, st); });
    if (!rc) m->g_ba.assign(MIXED_CHUNK_MAX, nullptr);
    if (rc) destroy_graphs(m);   // retried from scratch on the next call
    return rc;
}

// k PCG steps (b, a) in one graph
int run_ba(gpaw_poisson* ps, MixedState* m, int k, hipStream_t st)
{
    hipGraphExec_t& g = m->g_ba[k - 1];
    if (!g) {
        int rc = capture_graph(&g, [&](hipStream_t s) -> hipError_t {
            hipError_t e = hipSuccess;
            for (int i = 0; i < k && e == hipSuccess; i++) {
                e = emit_b(ps, m, s);
                if (e == hipSuccess) e = emit_a(ps, m, s, i == k - 1);
            }
            return e;
        });
        if (rc) return rc;
    }
    return gpaw_check(hipGraphLaunch(g, st), "mixed: iterations");
}

int read_scalar(MixedState* m, int slot, hipStream_t st, double* v)
{
    hipError_t e = hipMemcpyAsync(m->h_S, m->S, sizeof(double) * S_N, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return gpaw_check(e, "mixed: readback");
    *v = m->h_S[slot];
    return GPAW_OK;
}

// d = L phi - rho (FP64, stored in res[0]) with |d|^2 into S[DD]; d32 = d
int residual64(gpaw_poisson* ps, MixedState* m, hipStream_t st)
{
    gpaw_operator* op = ps->ops[0];
    Phases ph;
    gpaw_phases(nullptr, ph);
    int np = 0;
    hipError_t e = launch_stencil(op->st, DT_F64, ST_RESID, ps->phiA[0], ps->res[0], ps->rho[0],
                                  op->g, ph, 1, 0.0, op->partial, &np, true, st);
    if (e == hipSuccess) e = launch_reduce_partials(op->partial, np, 1.0, m->S + S_DD, st);
    if (e == hipSuccess) e = launch_cast(m->rho[0], DT_F32, ps->res[0], DT_F64, ps->npts[0], st);
    return gpaw_check(e, "mixed: residual");
}

}  // namespace

void gpaw_mixed_free(MixedState* m)
{
    if (!m) return;
    for (hipGraphExec_t g : {m->g_vcycle, m->g_vwarm, m->g_a, m->g_b})
        if (g) hipGraphExecDestroy(g);
    for (auto v : {&m->phiA, &m->phiB, &m->rho, &m->res})
        for (float* p : *v)
            if (p) hipFree(p);
    for (double* p : {m->p, m->q, m->S, m->partial})
        if (p) hipFree(p);
    if (m->h_S) hipHostFree(m->h_S);
    if (m->d_coarse32) hipFree(m->d_coarse32);
    delete m;
}

extern "C" int gpaw_poisson_solve_mixed(gpaw_poisson* ps, double* phi, const double* rho,
                                        double rel_tol, int maxiter, int mode, int inner,
                                        int* niter, double* relres, void* stream)
{
    if (!ps || !phi || !rho || maxiter < 1 || rel_tol <= 0.0 || (mode != 0 && mode != 1) ||
        inner < 1)
        return gpaw_fail(GPAW_EINVAL, "mixed: bad argument");
    hipStream_t st = (hipStream_t)stream;
    int rc = setup(ps);
    if (!rc && !ps->mixed->tuned && autotune_set(-1)) {
        // one eager pass of the graph bodies on the plan's own buffers (all
        // reloaded below), so that the graphs record tuned launch shapes
        ps->mixed->tuned = 1;
        hipError_t e = emit32_top(ps, ps->mixed, true, st);
        if (e == hipSuccess) e = emit_a(ps, ps->mixed, st);
        rc = gpaw_check(e, "mixed: tune");
    }
    if (!rc) rc = capture_all(ps);
    if (rc) return rc;
    MixedState* m = ps->mixed;
    const i64 n = ps->npts[0];
    hipError_t e = launch_copy(ps->rho[0], rho, DT_F64, n, st);
    if (e == hipSuccess) e = launch_copy(ps->phiA[0], phi, DT_F64, n, st);
    if (e == hipSuccess) {   // |rho|^2 for the relative criterion
        launch_kernel(k_dot, gridn(n), dim3(256), 0, st, ps->rho[0], ps->rho[0], n, m->partial);
        launch_kernel(k_red1, dim3(1), dim3(256), 0, st, m->partial, (int)gridn(n).x, m->S,
                           (int)S_RR);
        e = launch_error();
    }
    if (e != hipSuccess) return gpaw_check(e, "mixed: load");
    double rr = 0.0, dd = 0.0;
    // |rho|^2 and the first |d|^2 come back in one readback
    rc = residual64(ps, m, st);
    if (!rc) rc = read_scalar(m, S_DD, st, &dd);
    if (rc) return rc;
    rr = m->h_S[S_RR];
    const double target = rel_tol * rel_tol * rr;
    int it = 0;
    if (!rc && mode == 1 && dd > target) {
        // z = V32(d); dz = d.z; p = z
        rc = gpaw_check(hipGraphLaunch(m->g_vcycle, st), "mixed: V-cycle");
        if (!rc) {
            launch_kernel(k_dot_dz_zq, gridn(n), dim3(256), 0, st, ps->res[0], z_of(m),
                               ps->res[0], n, m->partial);
            launch_kernel(k_p_update, gridn(n), dim3(256), 0, st, m->p, z_of(m),
                          (const double*)m->S, n, 1, (const double*)nullptr, 0);
            rc = gpaw_check(launch_error(), "mixed: start");
        }
        // PCG steps run in chunks of k (b, a) between readbacks of |d|^2; a
        // chunk may run a few steps past the tolerance, which only makes the
        // result more accurate.  k follows the observed convergence rate.
        if (!rc) rc = gpaw_check(hipGraphLaunch(m->g_a, st), "mixed: iteration");
        if (!rc) rc = read_scalar(m, S_DD, st, &dd);
        it = 1;
        double prev = dd;
        int k = 2;
        while (!rc && dd > target && it < maxiter) {
            if (k > maxiter - it) k = maxiter - it;
            rc = run_ba(ps, m, k, st);
            if (!rc) rc = read_scalar(m, S_DD, st, &dd);
            it += k;
            // steps still needed at the rate of this chunk (|d|^2 per step)
            const int done = k;
            k = MIXED_CHUNK_MAX;
            if (!rc && dd > target && dd < prev && prev > 0) {
                const double rate = std::pow(dd / prev, 1.0 / done);
                if (rate > 0 && rate < 1) {
                    const double need = std::ceil(std::log(target / dd) / std::log(rate));
                    if (need >= 1 && need < MIXED_CHUNK_MAX) k = (int)need;
                }
            }
            prev = dd;
        }
    } else if (!rc && mode == 0) {
        while (!rc && dd > target && it < maxiter) {
            rc = gpaw_check(hipGraphLaunch(m->g_vcycle, st), "mixed: V-cycle");
            for (int k = 1; k < inner && !rc; k++)
                rc = gpaw_check(hipGraphLaunch(m->g_vwarm, st), "mixed: V-cycle");
            if (!rc) {
                launch_kernel(k_sub_f32, gridn(n), dim3(256), 0, st, ps->phiA[0], z_of(m), n);
                rc = gpaw_check(launch_error(), "mixed: correction");
            }
            if (!rc) rc = residual64(ps, m, st);
            if (!rc) rc = read_scalar(m, S_DD, st, &dd);
            it++;
        }
    }
    if (rc) return rc;
    rc = gpaw_check(launch_copy(phi, ps->phiA[0], DT_F64, n, st), "mixed: store");
    if (niter) *niter = it;
    if (relres) *relres = rr > 0 ? sqrt(dd / rr) : 0.0;
    if (rc) return rc;
    if (dd > target)
        return gpaw_fail(GPAW_ENOCONV, "mixed-precision Poisson solver did not reach %g in %d "
                                       "iterations", rel_tol, maxiter);
    return GPAW_OK;
}
