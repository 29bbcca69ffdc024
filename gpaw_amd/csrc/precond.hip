// Multigrid preconditioner of the finite-difference eigensolvers
// (gpaw/preconditioner.py:13-70) as one stream-ordered sequence of ten
// launches over all bands at once.
//
// The reference evaluates, per call (r = residuals, s = step):
//   r1 = R0(-r);  d1 = 4s r1;  q1 = L1 d1 - r1
//   r2 = R1(q1);  d2 = 16s r2; q2 = L2 d2 - r2;  d2 -= 16s q2
//   d1 -= I2(d2); q1 = L1 d1 - r1;  d1 -= 4s q1
//   d0 = I1(-d1); q0 = L0 d0 - r;   d0 += -s q0;  d0 *= -1
// Restriction, interpolation and the stencils are linear and round
// symmetrically, so R0(-r) == -R0(r) and I1(-d1) == -I1(d1) bit for bit.
// With e0 = I1(d1) = -d0 the last line becomes e0 - s (L0 e0 + r), which is
// one fused stencil pass (ST_PCFIN); the two "d -= w (L d - r)" updates are
// ST_PCUPD passes and "d1 -= I2(d2)" is the subtracting interpolation.
// Every value is rounded exactly where the reference's numpy / BLAS steps
// round it (BLAS axpy evaluated as y + (a x), without a fused multiply-add).
//
// Buffers (caller's workspace, one slice per band, 256-byte aligned):
//   level 0: E0;  level 1: R1, D1, Q1;  level 2: R2, D2, X2.
#include "capi_internal.h"

namespace gpaw {

// r = -r (if neg); d = c * r, over n doubles (complex data counted as pairs)
__global__ __launch_bounds__(256) void k_pc_scale(double* __restrict__ r, double* __restrict__ d,
                                                  double c, i64 n, int neg)
{
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double v = r[i];
    if (neg) {
        v = -v;
        r[i] = v;
    }
    d[i] = c * v;
}

}  // namespace gpaw

using namespace gpaw;

struct gpaw_preconditioner {
    gpaw_operator* kin[3];
    gpaw_transformer* rst[2];   // level 0 -> 1, 1 -> 2
    gpaw_transformer* ip[2];    // level 2 -> 1, 1 -> 0
    double step;
    i64 ng[3];
};

static i64 npts(const int n[3]) { return (i64)n[0] * n[1] * n[2]; }

static bool same(const int a[3], const int b[3])
{
    return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}

extern "C" int gpaw_preconditioner_new(gpaw_operator* kin0, gpaw_operator* kin1,
                                       gpaw_operator* kin2, gpaw_transformer* restrictor0,
                                       gpaw_transformer* restrictor1,
                                       gpaw_transformer* interpolator2,
                                       gpaw_transformer* interpolator1,
                                       gpaw_preconditioner** out)
{
    if (!out) return gpaw_fail(GPAW_EINVAL, "preconditioner: NULL argument");
    *out = nullptr;
    if (!kin0 || !kin1 || !kin2 || !restrictor0 || !restrictor1 || !interpolator2 ||
        !interpolator1)
        return gpaw_fail(GPAW_EINVAL, "preconditioner: NULL operator or transformer");
    const Transfer &r0 = restrictor0->t, &r1 = restrictor1->t;
    const Transfer &i2 = interpolator2->t, &i1 = interpolator1->t;
    if (r0.interpolate || r1.interpolate || !i2.interpolate || !i1.interpolate)
        return gpaw_fail(GPAW_EINVAL,
                         "preconditioner: expects two restrictions and two interpolations");
    if (!same(kin0->g.n, r0.gin.n) || !same(r0.nout, kin1->g.n) || !same(kin1->g.n, r1.gin.n) ||
        !same(r1.nout, kin2->g.n) || !same(kin2->g.n, i2.gin.n) || !same(i2.nout, kin1->g.n) ||
        !same(kin1->g.n, i1.gin.n) || !same(i1.nout, kin0->g.n))
        return gpaw_fail(GPAW_EINVAL, "preconditioner: grid sizes of the levels do not chain");
    const int real = kin0->real;
    if (kin1->real != real || kin2->real != real || restrictor0->real != real ||
        restrictor1->real != real || interpolator2->real != real || interpolator1->real != real)
        return gpaw_fail(GPAW_EINVAL, "preconditioner: mixed real and complex objects");
    if (kin0->diag == 0.0)
        return gpaw_fail(GPAW_EINVAL, "preconditioner: kin0 has no diagonal element");
    gpaw_preconditioner* pc = new gpaw_preconditioner();
    pc->kin[0] = kin0;
    pc->kin[1] = kin1;
    pc->kin[2] = kin2;
    pc->rst[0] = restrictor0;
    pc->rst[1] = restrictor1;
    pc->ip[0] = interpolator2;
    pc->ip[1] = interpolator1;
    pc->step = 0.66666666 / kin0->diag;   // gpaw/preconditioner.py:24
    for (int l = 0; l < 3; l++) pc->ng[l] = npts(pc->kin[l]->g.n);
    *out = pc;
    return GPAW_OK;
}

extern "C" double gpaw_preconditioner_step(const gpaw_preconditioner* pc)
{
    return pc ? pc->step : 0.0;
}

static size_t slice_bytes(i64 n, long nin, int dtype)
{
    const size_t b = (size_t)n * (size_t)nin * (size_t)dtype_size(dtype);
    return (b + 255) & ~(size_t)255;
}

static size_t workspace_bytes(const gpaw_preconditioner* pc, long nin, int dtype)
{
    return slice_bytes(pc->ng[0], nin, dtype) + 3 * slice_bytes(pc->ng[1], nin, dtype) +
           3 * slice_bytes(pc->ng[2], nin, dtype);
}

static int check_args(const gpaw_preconditioner* pc, long nin, int dtype)
{
    if (!pc) return gpaw_fail(GPAW_EINVAL, "preconditioner: NULL handle");
    if (dtype != GPAW_F64 && dtype != GPAW_C128)
        return gpaw_fail(GPAW_EUNSUPPORTED, "preconditioner: float64 or complex128 data only");
    int rc = gpaw_check_dtype(pc->kin[0]->real, dtype);
    if (rc) return rc;
    if (nin < 0 || nin > 65535) return gpaw_fail(GPAW_EINVAL, "preconditioner: nin = %ld", nin);
    return GPAW_OK;
}

extern "C" int gpaw_preconditioner_workspace(const gpaw_preconditioner* pc, long nin, int dtype,
                                             long* bytes)
{
    int rc = check_args(pc, nin, dtype);
    if (rc) return rc;
    if (!bytes) return gpaw_fail(GPAW_EINVAL, "preconditioner: NULL bytes");
    *bytes = (long)workspace_bytes(pc, nin, dtype);
    return GPAW_OK;
}

extern "C" int gpaw_preconditioner_apply(gpaw_preconditioner* pc, const void* residuals,
                                         void* out, long nin, int dtype, const double* phases,
                                         void* work, long work_bytes, void* stream)
{
    int rc = check_args(pc, nin, dtype);
    if (rc) return rc;
    if (nin == 0) return GPAW_OK;
    if (!residuals || !out || !work) return gpaw_fail(GPAW_EINVAL, "preconditioner: NULL array");
    if (residuals == out) return gpaw_fail(GPAW_EINVAL, "preconditioner: out must differ from residuals");
    if (work_bytes < 0 || (size_t)work_bytes < workspace_bytes(pc, nin, dtype))
        return gpaw_fail(GPAW_EINVAL, "preconditioner: workspace of %ld bytes, %zu needed",
                         work_bytes, workspace_bytes(pc, nin, dtype));
    if ((reinterpret_cast<uintptr_t>(work) & 255) != 0)
        return gpaw_fail(GPAW_EINVAL, "preconditioner: workspace must be 256-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    Phases ph;
    gpaw_phases(phases, ph);
    const int nb = (int)nin;
    char* w = (char*)work;
    auto take = [&](int l) {
        void* p = w;
        w += slice_bytes(pc->ng[l], nin, dtype);
        return p;
    };
    void* E0 = take(0);
    void* R1 = take(1);
    void* D1 = take(1);
    void* Q1 = take(1);
    void* R2 = take(2);
    void* D2 = take(2);
    void* X2 = take(2);
    const int dpp = dtype_complex(dtype) ? 2 : 1;   // doubles per point
    const i64 n1 = pc->ng[1] * nin * dpp, n2 = pc->ng[2] * nin * dpp;
    const double s = pc->step, c1 = 4 * s, c2 = 16 * s;
    gpaw_operator *L0 = pc->kin[0], *L1 = pc->kin[1], *L2 = pc->kin[2];

    hipError_t e = launch_transfer(pc->rst[0]->t, dtype, TR_STORE, residuals, R1, nullptr, ph, nb, st);
    if (e == hipSuccess) {   // r1 = -R0(r), d1 = 4s r1
        launch_kernel(k_pc_scale, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, st,
                           (double*)R1, (double*)D1, c1, n1, 1);
        e = launch_error();
    }
    if (e == hipSuccess)   // q1 = L1 d1 - r1
        e = launch_stencil(L1->st, dtype, ST_RESID, D1, Q1, R1, L1->g, ph, nb, 0.0, nullptr,
                           nullptr, true, st);
    if (e == hipSuccess) e = launch_transfer(pc->rst[1]->t, dtype, TR_STORE, Q1, R2, nullptr, ph, nb, st);
    if (e == hipSuccess) {   // d2 = 16s r2
        launch_kernel(k_pc_scale, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, st,
                           (double*)R2, (double*)D2, c2, n2, 0);
        e = launch_error();
    }
    if (e == hipSuccess)   // d2 - 16s (L2 d2 - r2)
        e = launch_stencil(L2->st, dtype, ST_PCUPD, D2, X2, R2, L2->g, ph, nb, c2, nullptr,
                           nullptr, true, st);
    if (e == hipSuccess)   // d1 -= I2(d2)
        e = launch_transfer(pc->ip[0]->t, dtype, TR_SUBTRACT, X2, D1, nullptr, ph, nb, st);
    if (e == hipSuccess)   // d1 - 4s (L1 d1 - r1)
        e = launch_stencil(L1->st, dtype, ST_PCUPD, D1, Q1, R1, L1->g, ph, nb, c1, nullptr,
                           nullptr, true, st);
    if (e == hipSuccess)   // e0 = I1(d1) = -d0
        e = launch_transfer(pc->ip[1]->t, dtype, TR_STORE, Q1, E0, nullptr, ph, nb, st);
    if (e == hipSuccess)   // e0 - s (L0 e0 + r)
        e = launch_stencil(L0->st, dtype, ST_PCFIN, E0, out, residuals, L0->g, ph, nb, s, nullptr,
                           nullptr, true, st);
    return gpaw_check(e, "preconditioner apply");
}

extern "C" void gpaw_preconditioner_free(gpaw_preconditioner* pc) { delete pc; }
