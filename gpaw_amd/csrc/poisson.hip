// Device-resident multigrid V-cycle (PoissonSolver.iterate2 /
// solve_neutral, gpaw/poisson.py:193-218,245-276).
//
// The plan owns every level array: rho_l, the residual r_l and a ping-pong
// pair for phi_l (weighted Jacobi reads the old field and writes a new one,
// c/bmgs/relax.c:47-71).  One V-cycle is a fixed sequence of kernels:
//
//   level l < L:  relax(pre_l) ; r_l = L_l phi_l - rho_l ;
//                 rho_{l+1} = R r_l and phi_{l+1} = 0 (one fused kernel) ;
//                 V-cycle(l+1) ; phi_l -= I phi_{l+1} (fused) ;
//   every level:  relax(post_l)
//   level 0:      dv * sum (L phi - rho)^2 into a device scalar (no r stored)
//
// identical in arithmetic to the reference's sequence of Operator /
// Transformer calls and numpy updates.  The sequence is captured once into
// a hipGraph and replayed; when level 0 does an odd number of sweeps per
// cycle the phi_0 buffers swap roles every cycle, so two graphs (one per
// starting buffer) are captured and alternated.  The host reads the 8-byte
// error after each cycle to apply the reference's stopping rule exactly.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "poisson_internal.h"

using namespace gpaw;


static int setup_coarse(gpaw_poisson* ps);

namespace {

int alloc(double** p, long long n)
{
    return gpaw_check(hipMalloc((void**)p, sizeof(double) * (n > 0 ? n : 1)), "poisson: level array");
}

// Whether level l's post-smoothing starts with a fused Jacobi launch that
// can apply the coarse-grid correction itself.
bool folds_correction(const gpaw_poisson* ps, int l)
{
    return ps->fold_correction && ps->method == 2 && l > 0 && l < ps->L && ps->fuse_jacobi &&
           jacobi_fused_ok(ps->ops[l]->st, ps->ops[l]->g) &&
           ps->post[l] >= jacobi_fused_sweeps(ps->ops[l]->g) &&
           ps->itp[l]->t.K == 2;
}

// Whether level 0's post-smoothing starts with a sweep that applies the
// coarse-grid correction itself (never the cycle's first sweep, which may
// sum the norm of the cycle's input).
bool folds_correction0(const gpaw_poisson* ps, int l)
{
    return ps->fold_correction && ps->method == 2 && l == 0 && l < ps->L && ps->pre[0] >= 1 &&
           ps->post[0] >= 1 && jacobi_corr_ok(ps->ops[0]->st, ps->ops[0]->g, ps->itp[0]->t);
}

// relax k sweeps on level l, ping-ponging between phi[l] and alt; returns via
// `cur` which buffer holds the result.  On the levels below 0, runs of
// 4 (2 on large levels) Jacobi sweeps are one fused launch (the result then sits in the
// other buffer than after as many single sweeps; callers follow `cur`).
// `corr` (with folds_correction): the first fused launch first applies
// phi -= I corr.  `zero_in`: phi is zero on entry (the first fused launch
// does not read it).
hipError_t relax(gpaw_poisson* ps, int l, int k, double*& cur, double*& alt, hipStream_t st,
                 const double* corr = nullptr, bool zero_in = false)
{
    Phases ph;
    gpaw_phases(nullptr, ph);
    gpaw_operator* op = ps->ops[l];
    if (ps->method == 1) {   // Gauss-Seidel: in place, alt is the halo snapshot
        if (k <= 0) return hipSuccess;
        return gpaw_op_gauss_seidel(op, cur, ps->rho[l], k, alt, st) == GPAW_OK
                   ? hipSuccess
                   : hipErrorLaunchFailure;
    }
    if (l == 0 && ps->rot.total) {   // chunked solve: the cycle's input is kept
        for (int it = 0; it < k; it++) {
            // the cycle's first sweep may also sum the previous cycle's error
            // (the residual norm of its input) into ps->norm_in
            double* part = ps->rot.done == 0 ? ps->norm_in : nullptr;
            double* out = ++ps->rot.done == ps->rot.total ? ps->rot.dst
                                                          : ps->rot.t[(ps->rot.done - 1) & 1];
            int np = 0;
            hipError_t e;
            if (corr) {   // the correction's sweep never sums the cycle's input norm
                e = part ? hipErrorInvalidValue
                         : launch_jacobi_corr(op->st, DT_F64, cur, out, ps->rho[l], corr, op->g,
                                              ps->w[l], ps->itp[l]->t, st);
                corr = nullptr;
            } else {
                e = launch_stencil(op->st, DT_F64, ST_JACOBI, cur, out, ps->rho[l], op->g, ph, 1,
                                   ps->w[l], part, part ? &np : nullptr, true, st);
            }
            if (e == hipSuccess && part && np != ps->np0) e = hipErrorInvalidValue;
            if (e != hipSuccess) return e;
            cur = out;
        }
        return hipSuccess;
    }
    if (l > 0 && ps->fuse_jacobi && jacobi_fused_ok(op->st, op->g)) {
        const int S = jacobi_fused_sweeps(op->g);
        for (; k >= S; k -= S) {
            hipError_t e = launch_jacobi_fused(op->st, DT_F64, cur, alt, ps->rho[l], op->g,
                                               ps->w[l], st, corr, corr ? &ps->itp[l]->t : nullptr,
                                               zero_in, S);
            corr = nullptr;
            zero_in = false;
            if (e != hipSuccess) return e;
            double* t = cur;
            cur = alt;
            alt = t;
        }
    }
    for (int it = 0; it < k; it++) {
        hipError_t e =
            corr ? launch_jacobi_corr(op->st, DT_F64, cur, alt, ps->rho[l], corr, op->g, ps->w[l],
                                      ps->itp[l]->t, st)
                 : launch_stencil(op->st, DT_F64, ST_JACOBI, cur, alt, ps->rho[l], op->g, ph, 1,
                                  ps->w[l], nullptr, nullptr, true, st);
        corr = nullptr;
        if (e != hipSuccess) return e;
        double* t = cur;
        cur = alt;
        alt = t;
    }
    return hipSuccess;
}

hipError_t emit_level(gpaw_poisson* ps, int l, double*& cur0, double*& alt0, hipStream_t st)
{
    Phases ph;
    gpaw_phases(nullptr, ph);
    if (l == ps->lc) {
        // the fused coarse kernel leaves its result where single sweeps would
        const int k = (l < ps->L ? ps->pre[l] : 0) + ps->post[l];
        cur0 = (ps->method == 2 && k % 2) ? ps->phiB[l] : ps->phiA[l];
        return launch_coarse_vcycle(ps->d_coarse, ps->h_coarse, ps->coarse_lds, st);
    }
    double* cur = l == 0 ? cur0 : ps->phiA[l];
    double* alt = l == 0 ? alt0 : ps->phiB[l];
    hipError_t e = hipSuccess;
    gpaw_operator* op = ps->ops[l];
    if (l < ps->L) {
        // below level 0, phi_l was zeroed by the restriction onto it
        e = relax(ps, l, ps->pre[l], cur, alt, st, nullptr, l > 0);
        if (e != hipSuccess) return e;
        // r_l = L_l phi_l - rho_l ; rho_{l+1} = R r_l ; phi_{l+1} = 0
        if (ps->fuse_restrict && resid_restrict_ok(op->st, op->g, ps->rst[l]->t)) {
            e = launch_resid_restrict(op->st, DT_F64, cur, ps->rho[l], ps->rho[l + 1],
                                      ps->phiA[l + 1], op->g, ps->rst[l]->t, st);
        } else {
            e = launch_stencil(op->st, DT_F64, ST_RESID, cur, ps->res[l], ps->rho[l], op->g, ph,
                               1, 0.0, nullptr, nullptr, true, st);
            if (e == hipSuccess)
                e = launch_transfer(ps->rst[l]->t, DT_F64, TR_STORE, ps->res[l], ps->rho[l + 1],
                                    ps->phiA[l + 1], ph, 1, st);
        }
        if (e != hipSuccess) return e;
        // the next level starts from phiA[l+1]; ccur returns its result buffer
        double* ccur = nullptr;
        double* calt = nullptr;
        e = emit_level(ps, l + 1, ccur, calt, st);
        if (e != hipSuccess) return e;
        if (folds_correction(ps, l)) {   // phi_l -= I phi_{l+1} inside the first sweeps
            e = relax(ps, l, ps->post[l], cur, alt, st, ccur);
            if (e != hipSuccess) return e;
            cur0 = cur;
            alt0 = alt;
            return e;
        }
        if (folds_correction0(ps, l)) {   // phi_0 -= I phi_1 inside the first sweep
            e = relax(ps, l, ps->post[l], cur, alt, st, ccur);
        } else {
            e = launch_transfer(ps->itp[l]->t, DT_F64, TR_SUBTRACT, ccur, cur, nullptr, ph, 1,
                                st);
            if (e == hipSuccess) e = relax(ps, l, ps->post[l], cur, alt, st);
        }
    } else {
        e = relax(ps, l, ps->post[l], cur, alt, st);
    }
    if (e != hipSuccess) return e;
    if (l == 0 && !ps->skip_norm) {
        int np = 0;
        double* part = ps->defer_partial ? ps->defer_partial : op->partial;
        e = launch_stencil(op->st, DT_F64, ST_RESID, cur, nullptr, ps->rho[0], op->g, ph, 1, 0.0,
                           part, &np, false, st);
        if (e != hipSuccess) return e;
        if (ps->defer_partial)   // summed with the chunk's other norms
            e = np == ps->np0 ? hipSuccess : hipErrorInvalidValue;
        else
            e = launch_reduce_partials(op->partial, np, ps->dv,
                                       ps->err_dst ? ps->err_dst : ps->d_err, st);
    }
    cur0 = cur;
    alt0 = alt;
    return e;
}

hipError_t emit_cycle(gpaw_poisson* ps, int start, hipStream_t st)
{
    double* cur = start ? ps->phiB[0] : ps->phiA[0];
    double* alt = start ? ps->phiA[0] : ps->phiB[0];
    return emit_level(ps, 0, cur, alt, st);
}

int capture(gpaw_poisson* ps)
{
    if (ps->exec[0]) return GPAW_OK;
    for (int s = 0; s < 1 + ps->swaps; s++) {
        hipError_t e = build_graph(&ps->exec[s],
                                   [&](hipStream_t st) { return emit_cycle(ps, s, st); });
        if (e != hipSuccess) {
            for (int t = 0; t < s; t++) {
                hipGraphExecDestroy(ps->exec[t]);
                ps->exec[t] = nullptr;
            }
            return gpaw_check(e, "poisson: V-cycle graph");
        }
    }
    return GPAW_OK;
}

// k V-cycles; the error after cycle i goes to d_errs[i].
// Jacobi: from phiA[0]; cycle i reads S_i and leaves its result in
// S_{i+1} = snap[i] (S_0 = phiA[0]) without writing S_i, its other sweeps
// going to phiB[0] and tmp2; then S_k is copied to phiA[0] for the next
// chunk.  Phi after every cycle of the chunk thus stays available with one
// copy per chunk.  Gauss-Seidel (in place): from `state`, phi copied to
// snap[i] before cycle i.
bool chunk_rotates(const gpaw_poisson* ps)
{
    // Jacobi with level 0 swept by single launches (not the fused coarse
    // kernel), and a first level-0 sweep that reads the cycle's input before
    // anything else touches it: with levels below 0 that is a pre-smoothing
    // sweep (without one, the coarse-grid correction would be subtracted in
    // place from the previous cycle's kept result, and a fused norm would be
    // summed from the corrected field); on a single level, a post-smoothing
    // sweep.  Other schedules copy phi before each cycle instead.
    if (ps->method != 2 || ps->lc == 0) return false;
    return ps->L > 0 ? ps->pre[0] > 0 : ps->post[0] > 0;
}

hipError_t emit_chunk(gpaw_poisson* ps, int k, int state, hipStream_t st)
{
    hipError_t e = hipSuccess;
    if (chunk_rotates(ps)) {
        const int n = (ps->L > 0 ? ps->pre[0] : 0) + ps->post[0];
        for (int i = 0; i < k && e == hipSuccess; i++) {
            double* cur = i ? ps->snap[i - 1] : ps->phiA[0];
            double* alt = ps->phiB[0];
            ps->rot.dst = ps->snap[i];
            ps->rot.t[0] = ps->phiB[0];
            ps->rot.t[1] = ps->tmp2;
            ps->rot.done = 0;
            ps->rot.total = n;
            ps->defer_partial = ps->d_chunk_partials + (long long)i * ps->np0;
            // with fused norms, cycle i's error is summed by cycle i + 1's
            // first sweep; the chunk's last cycle takes its own norm pass
            const bool fuse = ps->fuse_norm && stencil_dot_ok(ps->ops[0]->st, DT_F64,
                                                              ps->ops[0]->g, 1);
            ps->norm_in = (fuse && i > 0) ? ps->d_chunk_partials + (long long)(i - 1) * ps->np0
                                          : nullptr;
            ps->skip_norm = fuse && i < k - 1;
            e = emit_level(ps, 0, cur, alt, st);
            ps->defer_partial = nullptr;
            ps->norm_in = nullptr;
            ps->skip_norm = false;
            ps->rot.total = 0;
            if (e == hipSuccess && cur != ps->snap[i]) e = hipErrorInvalidValue;
        }
        // the k norms: row i of the partials, in the order of one reduction
        // launch per cycle
        if (e == hipSuccess)
            e = launch_reduce_rows(ps->d_chunk_partials, k, ps->np0, ps->dv, ps->d_errs, st);
        if (e == hipSuccess) e = launch_copy(ps->phiA[0], ps->snap[k - 1], DT_F64, ps->npts[0], st);
        return e;
    }
    for (int i = 0; i < k && e == hipSuccess; i++) {
        e = launch_copy(ps->snap[i], state ? ps->phiB[0] : ps->phiA[0], DT_F64, ps->npts[0], st);
        ps->err_dst = ps->d_errs + i;
        if (e == hipSuccess) e = emit_cycle(ps, state, st);
        ps->err_dst = nullptr;
        state ^= ps->swaps;
    }
    return e;
}

int ensure_chunks(gpaw_poisson* ps)
{
    if (!ps->snap.empty()) return GPAW_OK;
    int rc = GPAW_OK;
    ps->snap.assign(GPAW_CHUNK_MAX, nullptr);
    for (int i = 0; i < GPAW_CHUNK_MAX && !rc; i++) rc = alloc(&ps->snap[i], ps->npts[0]);
    if (!rc && chunk_rotates(ps)) {
        ps->np0 = stencil_partials(ps->ops[0]->st, ps->ops[0]->g, 1);
        rc = alloc(&ps->tmp2, ps->npts[0]);
        if (!rc) rc = alloc(&ps->d_chunk_partials, (long long)GPAW_CHUNK_MAX * ps->np0);
    }
    if (!rc) rc = alloc(&ps->d_errs, GPAW_CHUNK_MAX);
    if (!rc) rc = alloc(&ps->d_sum, 1024);
    if (!rc)
        rc = gpaw_check(hipHostMalloc((void**)&ps->h_errs, sizeof(double) * GPAW_CHUNK_MAX),
                        "poisson: pinned errors");
    if (!rc) ps->chunk.assign(2 * GPAW_CHUNK_MAX, nullptr);
    if (rc) {   // leave no half-built state behind
        for (double* p : ps->snap)
            if (p) hipFree(p);
        ps->snap.clear();
        if (ps->tmp2) hipFree(ps->tmp2);
        ps->tmp2 = nullptr;
        if (ps->d_chunk_partials) hipFree(ps->d_chunk_partials);
        ps->d_chunk_partials = nullptr;
        if (ps->d_errs) hipFree(ps->d_errs);
        ps->d_errs = nullptr;
        if (ps->d_sum) hipFree(ps->d_sum);
        ps->d_sum = nullptr;
    }
    return rc;
}

int run_chunk(gpaw_poisson* ps, int k, int state, hipStream_t st)
{
    if (ps->use_graph) {
        hipGraphExec_t& g = ps->chunk[2 * (k - 1) + state];
        if (!g) {
            hipError_t e = build_graph(&g, [&](hipStream_t s) { return emit_chunk(ps, k, state, s); });
            if (e != hipSuccess) return gpaw_check(e, "poisson: chunk graph");
        }
        return gpaw_check(hipGraphLaunch(g, st), "poisson: chunk launch");
    }
    return gpaw_check(emit_chunk(ps, k, state, st), "poisson: chunk");
}

int run_cycle(gpaw_poisson* ps, int state, hipStream_t st)
{
    if (ps->use_graph) {
        int rc = capture(ps);
        if (rc) return rc;
        return gpaw_check(hipGraphLaunch(ps->exec[state], st), "poisson: graph launch");
    }
    return gpaw_check(emit_cycle(ps, state, st), "poisson: V-cycle");
}

int read_error(gpaw_poisson* ps, hipStream_t st, double* err)
{
    hipError_t e = hipMemcpyAsync(ps->h_err, ps->d_err, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return gpaw_check(e, "poisson: error readback");
    *err = ps->h_err[0];
    return GPAW_OK;
}

}  // namespace

extern "C" int gpaw_poisson_new(int levels, gpaw_operator* const* ops,
                                gpaw_transformer* const* restrictors,
                                gpaw_transformer* const* interpolators, const int* presmooths,
                                const int* postsmooths, const double* weights, int relax_method,
                                double dv, gpaw_poisson** out)
{
    if (!out || !ops || !presmooths || !postsmooths || !weights || levels < 0 || levels > 16)
        return gpaw_fail(GPAW_EINVAL, "poisson: bad argument");
    *out = nullptr;
    if (relax_method != 1 && relax_method != 2)
        return gpaw_fail(GPAW_EINVAL, "poisson: unknown relaxation method %d", relax_method);
    if (levels > 0 && (!restrictors || !interpolators))
        return gpaw_fail(GPAW_EINVAL, "poisson: missing transformers");
    for (int l = 0; l <= levels; l++) {
        if (!ops[l] || !ops[l]->real) return gpaw_fail(GPAW_EINVAL, "poisson: operator %d", l);
        if (presmooths[l] < 0 || postsmooths[l] < 0)
            return gpaw_fail(GPAW_EINVAL, "poisson: negative smoothing count");
    }
    for (int l = 0; l < levels; l++) {
        const Transfer& r = restrictors[l]->t;
        const Transfer& i = interpolators[l]->t;
        for (int a = 0; a < 3; a++) {
            if (r.interpolate || !i.interpolate || r.gin.n[a] != ops[l]->g.n[a] ||
                r.nout[a] != ops[l + 1]->g.n[a] || i.gin.n[a] != ops[l + 1]->g.n[a] ||
                i.nout[a] != ops[l]->g.n[a])
                return gpaw_fail(GPAW_EINVAL, "poisson: level %d transformers do not match "
                                              "the operators", l);
        }
    }
    gpaw_poisson* ps = new gpaw_poisson();
    ps->L = levels;
    ps->method = relax_method;
    ps->dv = dv;
    ps->ops.assign(ops, ops + levels + 1);
    if (levels) {
        ps->rst.assign(restrictors, restrictors + levels);
        ps->itp.assign(interpolators, interpolators + levels);
    }
    ps->pre.assign(presmooths, presmooths + levels + 1);
    ps->post.assign(postsmooths, postsmooths + levels + 1);
    ps->w.assign(weights, weights + levels + 1);
    ps->swaps = relax_method == 2 ? ((levels > 0 ? ps->pre[0] : 0) + ps->post[0]) % 2 : 0;
    const char* env = getenv("GPAW_AMD_GRAPH");
    ps->use_graph = !(env && env[0] == '0');
    env = getenv("GPAW_AMD_FUSED_JACOBI");
    ps->fuse_jacobi = !(env && env[0] == '0');
    env = getenv("GPAW_AMD_FOLD_CORRECTION");
    ps->fold_correction = !(env && env[0] == '0');
    env = getenv("GPAW_AMD_FUSED_RESTRICT");
    ps->fuse_restrict = !(env && env[0] == '0');
    env = getenv("GPAW_AMD_FUSED_NORM");
    ps->fuse_norm = !(env && env[0] == '0');
    int rc = GPAW_OK;
    for (int l = 0; l <= levels && !rc; l++) {
        const Geom& g = ps->ops[l]->g;
        const long long n = (long long)g.n[0] * g.n[1] * g.n[2];
        ps->npts.push_back(n);
        double *a = nullptr, *b = nullptr, *r = nullptr, *x = nullptr;
        rc = alloc(&a, n);
        if (!rc) rc = alloc(&b, n);
        if (!rc) rc = alloc(&r, n);
        if (!rc) rc = alloc(&x, n);
        ps->phiA.push_back(a);
        ps->phiB.push_back(b);
        ps->rho.push_back(r);
        ps->res.push_back(x);
    }
    if (!rc) rc = gpaw_op_ensure_partials(ps->ops[0]);
    if (!rc && relax_method == 2) rc = setup_coarse(ps);
    for (int l = 0; l <= levels && !rc && relax_method == 1; l++) {
        gpaw_operator* op = ps->ops[l];
        if (!op->st.d_coefs) {   // Gauss-Seidel reads the stencil table on the device
            hipError_t e = hipMalloc(&op->st.d_coefs, sizeof(double) * op->st.nc);
            if (e == hipSuccess) e = hipMalloc(&op->st.d_off, sizeof(int) * 3 * op->st.nc);
            if (e == hipSuccess)
                e = hipMemcpy(op->st.d_coefs, op->st.coefs, sizeof(double) * op->st.nc,
                              hipMemcpyHostToDevice);
            if (e == hipSuccess)
                e = hipMemcpy(op->st.d_off, op->st.off, sizeof(int) * 3 * op->st.nc,
                              hipMemcpyHostToDevice);
            rc = gpaw_check(e, "poisson: stencil upload");
        }
    }
    if (!rc) rc = alloc(&ps->d_err, 1);
    if (!rc) rc = gpaw_check(hipHostMalloc((void**)&ps->h_err, sizeof(double)), "poisson: pinned");
    if (rc) {
        gpaw_poisson_free(ps);
        return rc;
    }
    *out = ps;
    return GPAW_OK;
}

// Levels from lc down are run by one fused kernel when they are small and
// all 7-point stars with K=2 transfers (the Poisson hierarchy below level 0).
static int setup_coarse(gpaw_poisson* ps)
{
    // the fused kernel holds at most 2048 points per level (two per thread)
    const char* env = getenv("GPAW_AMD_COARSE_MAX");
    long long limit = env ? atoll(env) : 2000;
    if (limit > 2048) limit = 2048;
    int lc = -1;
    for (int l = 1; l <= ps->L; l++)
        if (ps->npts[l] <= limit) {
            lc = l;
            break;
        }
    if (lc < 0 || ps->L - lc + 1 > GPAW_MAX_LEVELS) return GPAW_OK;
    CoarsePlan P;
    memset(&P, 0, sizeof P);
    P.nlev = ps->L - lc + 1;
    for (int l = lc; l <= ps->L; l++) {
        const gpaw_operator* op = ps->ops[l];
        if (op->st.star_r != 1) return GPAW_OK;
        if (l < ps->L && (ps->rst[l]->t.K != 2 || ps->itp[l]->t.K != 2)) return GPAW_OK;
        CoarseLevel& C = P.lev[l - lc];
        C.g = op->g;
        for (int c = 0; c < 7; c++) C.c[c] = op->st.coefs[c];
        C.w = ps->w[l];
        C.pre = l < ps->L ? ps->pre[l] : 0;
        C.post = ps->post[l];
        C.phiA = ps->phiA[l];
        C.phiB = ps->phiB[l];
        C.rho = ps->rho[l];
        C.res = ps->res[l];
        if (l < ps->L) {
            C.rst = make_trargs(ps->rst[l]->t);
            C.itp = make_trargs(ps->itp[l]->t);
        }
    }
    // neighbour tables (6 shorts per point), built here once
    std::vector<short> tab;
    for (int l = lc; l <= ps->L; l++) {
        P.lev[l - lc].nb_off = (int)tab.size();
        tab.resize(tab.size() + 6 * ps->npts[l]);
        coarse_neighbour_table(P.lev[l - lc].g, tab.data() + P.lev[l - lc].nb_off);
    }
    if (tab.size() % 2) tab.push_back(0);   // whole 4-byte words
    P.nb_words = (int)(tab.size() / 2);
    {
        hipError_t e = hipMalloc((void**)&ps->d_coarse_nb, sizeof(short) * tab.size());
        if (e == hipSuccess)
            e = hipMemcpy(ps->d_coarse_nb, tab.data(), sizeof(short) * tab.size(),
                          hipMemcpyHostToDevice);
        if (e != hipSuccess) return gpaw_check(e, "poisson: coarse neighbour tables");
        P.nbtab = ps->d_coarse_nb;
    }
    // keep the coarse arrays (4 per level, each with one more element: the
    // zero slot of the phi arrays) and the tables in LDS when they fit
    long long nd = 0;
    for (int l = lc; l <= ps->L; l++) {
        P.lev[l - lc].off = (int)nd;
        nd += 4 * (ps->npts[l] + 1);
    }
    if (nd * 8 + 4LL * P.nb_words <= 128 * 1024) {
        P.lds = 1;
        P.lds_doubles = (int)nd;
        ps->coarse_lds = (int)(nd * 8 + 4LL * P.nb_words);
        if (ps->coarse_lds > 64 * 1024) {
            hipError_t e = coarse_set_lds_limit(ps->coarse_lds);
            if (e != hipSuccess) {
                P.lds = 0;
                ps->coarse_lds = 0;
            }
        }
    }
    ps->h_coarse = P;
    hipError_t e = hipMalloc((void**)&ps->d_coarse, sizeof(CoarsePlan));
    if (e == hipSuccess) e = hipMemcpy(ps->d_coarse, &P, sizeof P, hipMemcpyHostToDevice);
    if (e != hipSuccess) return gpaw_check(e, "poisson: coarse plan");
    ps->lc = lc;
    return GPAW_OK;
}

// With the launch-shape tuner on, one V-cycle is first run eagerly on the
// plan's own buffers (their contents are overwritten by load_inputs), so
// that the graphs record tuned shapes: launches inside a graph are never
// timed.
static int pretune(gpaw_poisson* ps, hipStream_t st)
{
    if (ps->tuned || !autotune_set(-1)) return GPAW_OK;
    ps->tuned = 1;
    hipError_t e = hipSuccess;
    for (int l = 0; l <= ps->L && e == hipSuccess; l++) {
        e = launch_fill(ps->phiA[l], DT_F64, ps->npts[l], st);
        if (e == hipSuccess) e = launch_fill(ps->phiB[l], DT_F64, ps->npts[l], st);
        if (e == hipSuccess) e = launch_fill(ps->rho[l], DT_F64, ps->npts[l], st);
    }
    if (e == hipSuccess) e = emit_cycle(ps, 0, st);
    return gpaw_check(e, "poisson: tune");
}

static int load_inputs(gpaw_poisson* ps, const double* phi, const double* rho, hipStream_t st)
{
    int rc = pretune(ps, st);
    if (rc) return rc;
    ps->rho_loaded = false;
    hipError_t e = launch_copy(ps->rho[0], rho, DT_F64, ps->npts[0], st);
    if (e == hipSuccess) e = launch_copy(ps->phiA[0], phi, DT_F64, ps->npts[0], st);
    return gpaw_check(e, "poisson: load");
}

extern "C" int gpaw_poisson_vcycles(gpaw_poisson* ps, double* phi, const double* rho, int ncycles,
                                    double* errors, void* stream)
{
    if (!ps || !phi || !rho || ncycles < 0) return gpaw_fail(GPAW_EINVAL, "poisson: bad argument");
    hipStream_t st = (hipStream_t)stream;
    int rc = load_inputs(ps, phi, rho, st);
    int state = 0;
    for (int c = 0; c < ncycles && !rc; c++) {
        rc = run_cycle(ps, state, st);
        state ^= ps->swaps;
        double err = 0.0;
        if (!rc) rc = read_error(ps, st, &err);
        if (!rc && errors) errors[c] = err;
    }
    if (!rc)
        rc = gpaw_check(launch_copy(phi, state ? ps->phiB[0] : ps->phiA[0], DT_F64, ps->npts[0], st),
                        "poisson: store");
    return rc;
}

static int solve_loaded(gpaw_poisson* ps, double* phi, double eps, int maxiter, int* niter,
                        double* error, hipStream_t st);

extern "C" int gpaw_poisson_solve_neutral(gpaw_poisson* ps, double* phi, const double* rho,
                                          double eps, int maxiter, int* niter, double* error,
                                          void* stream)
{
    if (!ps || !phi || !rho || maxiter < 1) return gpaw_fail(GPAW_EINVAL, "poisson: bad argument");
    hipStream_t st = (hipStream_t)stream;
    int rc = load_inputs(ps, phi, rho, st);
    if (rc) return rc;
    return solve_loaded(ps, phi, eps, maxiter, niter, error, st);
}

extern "C" int gpaw_poisson_neutralize(gpaw_poisson* ps, const double* rho, double maxcharge,
                                       double* charge, int* neutral, void* stream)
{
    if (!ps || !rho || !charge || !neutral) return gpaw_fail(GPAW_EINVAL, "poisson: bad argument");
    hipStream_t st = (hipStream_t)stream;
    *neutral = 0;
    ps->rho_loaded = false;
    int rc = ensure_chunks(ps);
    if (rc) return rc;
    // the charge integral dv * sum(rho) of gd.integrate, the same kernels
    // and summation order as gpaw_sum
    const long long n = ps->npts[0];
    int np = 0;
    hipError_t e = launch_sum(rho, n, ps->d_sum, &np, st);
    if (e == hipSuccess) e = launch_reduce_partials(ps->d_sum, np, ps->dv, ps->d_err, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ps->h_err, ps->d_err, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return gpaw_check(e, "poisson: charge");
    const double q = ps->h_err[0];
    *charge = q;
    if (!(std::fabs(q) <= maxcharge)) return GPAW_OK;   // charged: the caller's branches
    // neutral: rho minus its mean, as gpaw/poisson.py:130-138 computes it,
    // written straight into the plan's level-0 rho (the eager pre-tuning
    // V-cycle, which works on the plan's arrays, runs first)
    rc = pretune(ps, st);
    if (rc) return rc;
    const double background = q / ps->dv / (double)n;
    e = launch_sub_scalar(ps->rho[0], rho, background, n, st);
    if (e != hipSuccess) return gpaw_check(e, "poisson: load");
    ps->rho_loaded = true;
    *neutral = 1;
    return GPAW_OK;
}

extern "C" int gpaw_poisson_solve(gpaw_poisson* ps, double* phi, const double* rho, double eps,
                                  int maxiter, double maxcharge, double* charge, int* neutral,
                                  int* niter, double* error, void* stream)
{
    if (!ps || !phi || !rho || !charge || !neutral || maxiter < 1)
        return gpaw_fail(GPAW_EINVAL, "poisson: bad argument");
    int rc = gpaw_poisson_neutralize(ps, rho, maxcharge, charge, neutral, stream);
    if (rc || !*neutral) return rc;
    ps->rho_loaded = false;
    hipStream_t st = (hipStream_t)stream;
    rc = gpaw_check(launch_copy(ps->phiA[0], phi, DT_F64, ps->npts[0], st), "poisson: load");
    if (rc) return rc;
    return solve_loaded(ps, phi, eps, maxiter, niter, error, st);
}

static int solve_loaded(gpaw_poisson* ps, double* phi, double eps, int maxiter, int* niter,
                        double* error, hipStream_t st)
{
    int rc = ensure_chunks(ps);
    if (rc) return rc;
    // The reference runs `while iterate2() > eps and niter < maxiter:
    // niter += 1` (gpaw/poisson.py:203-209), one V-cycle and one norm per
    // step.  Here V-cycles run in chunks of k with phi kept after each (see
    // emit_chunk); the host reads the k norms once, finds the first cycle
    // that meets the stopping rule and takes phi as it was after that
    // cycle.  niter and phi are exactly the reference's.  k follows the
    // observed convergence rate so that a chunk usually ends at the
    // stopping cycle; the first chunk is as long as the previous solve on
    // this plan took (repeated solves of similar densities, as in an SCF
    // loop, then read the norms back once).  Chunk lengths never change
    // niter or phi, only how often the host reads the norms.
    const bool jac = chunk_rotates(ps);
    const double* result = nullptr;
    int state = 0;
    int n = 0;         // V-cycles accepted so far
    double err = 0.0;
    double prev = -1.0, rate = -1.0;
    int k = (ps->hint > 0 && ps->hint_eps == eps) ? std::min(ps->hint, GPAW_CHUNK_MAX) : 2;
    for (;;) {
        if (k > maxiter - n) k = maxiter - n;
        rc = run_chunk(ps, k, state, st);
        if (rc) return rc;
        hipError_t e = hipMemcpyAsync(ps->h_errs, ps->d_errs, sizeof(double) * k,
                                      hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return gpaw_check(e, "poisson: error readback");
        int stop = -1;
        for (int i = 0; i < k; i++) {
            err = ps->h_errs[i];
            n++;
            if (prev > 0 && err > 0) rate = err / prev;
            prev = err;
            if (!(err > eps) || n >= maxiter) {
                stop = i;
                break;
            }
        }
        if (stop >= 0 && jac) {
            result = ps->snap[stop];
            break;
        }
        if (stop >= 0) {
            const int after = state ^ (ps->swaps & ((stop + 1) & 1));
            if (stop + 1 < k)   // cycles past the stopping one ran: take phi back
                e = launch_copy(after ? ps->phiB[0] : ps->phiA[0], ps->snap[stop + 1], DT_F64,
                                ps->npts[0], st);
            if (e != hipSuccess) return gpaw_check(e, "poisson: restore");
            state = after;
            break;
        }
        if (!jac) state ^= ps->swaps & (k & 1);
        // cycles still needed at the observed rate
        k = GPAW_CHUNK_MAX;
        if (rate > 0 && rate < 1 && err > eps) {
            const double m = std::ceil(std::log(eps / err) / std::log(rate));
            if (m >= 1 && m < GPAW_CHUNK_MAX) k = (int)m;
        }
    }
    if (!result) result = state ? ps->phiB[0] : ps->phiA[0];
    rc = gpaw_check(launch_copy(phi, result, DT_F64, ps->npts[0], st), "poisson: store");
    if (niter) *niter = n;
    if (error) *error = err;
    if (rc) return rc;
    if (n == maxiter)
        return gpaw_fail(GPAW_ENOCONV, "Poisson solver did not converge in %d iterations!",
                         maxiter);
    ps->hint = n;
    ps->hint_eps = eps;
    return GPAW_OK;
}

extern "C" void gpaw_poisson_free(gpaw_poisson* ps)
{
    if (!ps) return;
    for (int s = 0; s < 2; s++)
        if (ps->exec[s]) hipGraphExecDestroy(ps->exec[s]);
    for (auto v : {&ps->phiA, &ps->phiB, &ps->rho, &ps->res})
        for (double* p : *v)
            if (p) hipFree(p);
    if (ps->d_err) hipFree(ps->d_err);
    for (hipGraphExec_t g : ps->chunk)
        if (g) hipGraphExecDestroy(g);
    for (double* p : ps->snap)
        if (p) hipFree(p);
    if (ps->tmp2) hipFree(ps->tmp2);
    if (ps->d_chunk_partials) hipFree(ps->d_chunk_partials);
    if (ps->d_errs) hipFree(ps->d_errs);
    if (ps->d_sum) hipFree(ps->d_sum);
    if (ps->h_errs) hipHostFree(ps->h_errs);
    if (ps->d_coarse) hipFree(ps->d_coarse);
    if (ps->d_coarse_nb) hipFree(ps->d_coarse_nb);
    if (ps->mixed) gpaw_mixed_free(ps->mixed);
    if (ps->h_err) hipHostFree(ps->h_err);
    delete ps;
}
