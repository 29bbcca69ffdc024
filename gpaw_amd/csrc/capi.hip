// C ABI: Operator and Transformer objects (see include/gpaw_amd.h).
//
// An Operator holds the stencil (coefficients in the caller's order, the
// summation order of c/bmgs/fd.c:41-44) and the halo description; a
// Transformer holds the restriction / interpolation geometry.  Both are
// plans only: no grid-sized buffers are allocated except the relax
// ping-pong scratch and the norm partials, allocated on first use.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "capi_internal.h"

using namespace gpaw;

static thread_local std::string g_err;

int gpaw_fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int gpaw_check(hipError_t e, const char* what)
{
    if (e == hipSuccess) return GPAW_OK;
    return gpaw_fail(e == hipErrorOutOfMemory ? GPAW_ENOMEM : GPAW_EDEVICE, "%s: %s", what,
                     hipGetErrorString(e));
}

extern "C" const char* gpaw_last_error(void) { return g_err.c_str(); }

void gpaw_phases(const double* p, Phases& ph)
{
    for (int a = 0; a < 3; a++)
        for (int d = 0; d < 2; d++) {
            ph.re[a][d] = p ? p[4 * a + 2 * d] : 1.0;
            ph.im[a][d] = p ? p[4 * a + 2 * d + 1] : 0.0;
        }
}

static int neighbors_to_periodic(const long* nb, int per[3])
{
    for (int a = 0; a < 3; a++) {
        const long l = nb[2 * a], h = nb[2 * a + 1];
        if ((l != 0 && l != -1) || (h != 0 && h != -1))
            return gpaw_fail(GPAW_EUNSUPPORTED,
                             "neighbor rank %ld/%ld on axis %d: only a single domain "
                             "(0 = periodic, -1 = zero boundary) is supported",
                             l, h, a);
        if (l != h)
            return gpaw_fail(GPAW_EUNSUPPORTED,
                             "axis %d is periodic on one side only", a);
        per[a] = (l == 0);
    }
    return GPAW_OK;
}

// canonical orthogonal-Laplacian order: centre, then for axes 2, 1, 0:
// +1..+R, -1..-R (the order NewGUCLaplace produces, gpaw/fd_operators.py:266-275)
static int star_range(const Stencil& s)
{
    if ((s.nc - 1) % 6 != 0) return 0;
    const int R = (s.nc - 1) / 6;
    if (R < 1 || R > GPAW_MAXR) return 0;
    if (s.off[0][0] || s.off[0][1] || s.off[0][2]) return 0;
    int c = 1;
    for (int axis = 2; axis >= 0; axis--)
        for (int sgn = 1; sgn >= -1; sgn -= 2)
            for (int k = 1; k <= R; k++, c++)
                for (int a = 0; a < 3; a++)
                    if (s.off[c][a] != (a == axis ? sgn * k : 0)) return 0;
    return R;
}

extern "C" int gpaw_operator_new(const double* coefs, const long* offsets, int ncoefs,
                                 const long size[3], int range, const long* neighbors, int real,
                                 int cfd, gpaw_operator** out)
{
    if (!out || !coefs || !offsets || !size || !neighbors)
        return gpaw_fail(GPAW_EINVAL, "operator: NULL argument");
    *out = nullptr;
    if (ncoefs < 1 || ncoefs > GPAW_MAXC)
        return gpaw_fail(GPAW_EINVAL, "operator: ncoefs %d outside 1..%d", ncoefs, GPAW_MAXC);
    if (range < 0) return gpaw_fail(GPAW_EINVAL, "operator: negative range");
    for (int a = 0; a < 3; a++)
        if (size[a] < 1 || size[a] > (1L << 20))
            return gpaw_fail(GPAW_EINVAL, "operator: size[%d] = %ld", a, size[a]);
    gpaw_operator* op = new gpaw_operator();
    Stencil& s = op->st;
    s.nc = ncoefs;
    s.range = range;
    // decode the flat offsets of the padded buffer (extent size + 2*range)
    const long M1 = size[1] + 2 * range, M2 = size[2] + 2 * range;
    const long shift = range * (M1 * M2 + M2 + 1);
    for (int c = 0; c < ncoefs; c++) {
        const long o = offsets[c] + shift;
        const long d0 = o / (M1 * M2), r = o % (M1 * M2);
        const long d1 = r / M2, d2 = r % M2;
        if (o < 0 || d0 > 2 * range || d1 > 2 * range || d2 > 2 * range) {
            delete op;
            return gpaw_fail(GPAW_EINVAL, "operator: offset %ld outside the range %d box",
                             offsets[c], range);
        }
        s.off[c][0] = (int)(d0 - range);
        s.off[c][1] = (int)(d1 - range);
        s.off[c][2] = (int)(d2 - range);
        s.coefs[c] = coefs[c];
    }
    int rc = neighbors_to_periodic(neighbors, op->g.per);
    if (rc) {
        delete op;
        return rc;
    }
    for (int a = 0; a < 3; a++) {
        op->g.n[a] = (int)size[a];
        if (op->g.per[a] && range > size[a]) {
            delete op;
            return gpaw_fail(GPAW_EUNSUPPORTED,
                             "operator: periodic axis %d of %ld points narrower than the "
                             "stencil range %d",
                             a, size[a], range);
        }
    }
    op->real = real ? 1 : 0;
    op->cfd = cfd;
    op->diag = 0.0;
    for (int c = 0; c < ncoefs; c++)
        if (offsets[c] == 0) op->diag = coefs[c];
    s.star_r = star_range(s);
    hipGetDevice(&op->device);
    if (!s.star_r) {   // device copy of the stencil for the generic kernel
        hipError_t e = hipMalloc(&s.d_coefs, sizeof(double) * ncoefs);
        if (e == hipSuccess) e = hipMalloc(&s.d_off, sizeof(int) * 3 * ncoefs);
        if (e == hipSuccess)
            e = hipMemcpy(s.d_coefs, s.coefs, sizeof(double) * ncoefs, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(s.d_off, s.off, sizeof(int) * 3 * ncoefs, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            gpaw_operator_free(op);
            return gpaw_check(e, "operator: coefficient upload");
        }
    }
    *out = op;
    return GPAW_OK;
}

int gpaw_check_dtype(int real, int dtype)
{
    if (dtype < GPAW_F64 || dtype > GPAW_C64) return gpaw_fail(GPAW_EINVAL, "unknown dtype %d", dtype);
    if (real && dtype_complex(dtype))
        return gpaw_fail(GPAW_EINVAL, "complex data given to a real object");
    if (!real && !dtype_complex(dtype))
        return gpaw_fail(GPAW_EINVAL, "real data given to a complex object");
    return GPAW_OK;
}

int gpaw_op_ensure_partials(gpaw_operator* op)
{
    const int n = stencil_partials(op->st, op->g, 1);
    if (op->partial_n >= n) return GPAW_OK;
    if (op->partial) hipFree(op->partial);
    op->partial = nullptr;
    op->partial_n = 0;
    hipError_t e = hipMalloc(&op->partial, sizeof(double) * n);
    if (e != hipSuccess) return gpaw_check(e, "operator: partial sums");
    op->partial_n = n;
    return GPAW_OK;
}

int gpaw_op_ensure_scratch(gpaw_operator* op)
{
    if (op->scratch) return GPAW_OK;
    const size_t n = (size_t)op->g.n[0] * op->g.n[1] * op->g.n[2];
    return gpaw_check(hipMalloc(&op->scratch, sizeof(double) * n), "operator: relax scratch");
}

extern "C" int gpaw_operator_apply(gpaw_operator* op, const void* in, void* out, long nin,
                                   int dtype, const double* phases, void* stream)
{
    if (!op) return gpaw_fail(GPAW_EINVAL, "operator: NULL handle");
    int rc = gpaw_check_dtype(op->real, dtype);
    if (rc) return rc;
    if (nin < 0 || nin > 65535) return gpaw_fail(GPAW_EINVAL, "operator: nin = %ld", nin);
    if (nin == 0) return GPAW_OK;
    if (!in || !out) return gpaw_fail(GPAW_EINVAL, "operator: NULL array");
    if (in == out) return gpaw_fail(GPAW_EINVAL, "operator: in and out must differ");
    Phases ph;
    gpaw_phases(phases, ph);
    return gpaw_check(launch_stencil(op->st, dtype, ST_APPLY, in, out, nullptr, op->g, ph,
                                     (int)nin, 0.0, nullptr, nullptr, true,
                                     (hipStream_t)stream),
                      "operator apply");
}

extern "C" int gpaw_operator_apply_potential(gpaw_operator* op, const void* in, void* out,
                                             long nin, int dtype, const double* phases,
                                             const double* v, void* stream)
{
    if (!op) return gpaw_fail(GPAW_EINVAL, "operator: NULL handle");
    int rc = gpaw_check_dtype(op->real, dtype);
    if (rc) return rc;
    if (dtype != DT_F64 && dtype != DT_C128)
        return gpaw_fail(GPAW_EINVAL, "apply_potential: float64 or complex128 bands");
    if (nin < 0 || nin > 65535) return gpaw_fail(GPAW_EINVAL, "operator: nin = %ld", nin);
    if (nin == 0) return GPAW_OK;
    if (!in || !out || !v) return gpaw_fail(GPAW_EINVAL, "operator: NULL array");
    if (in == out) return gpaw_fail(GPAW_EINVAL, "operator: in and out must differ");
    Phases ph;
    gpaw_phases(phases, ph);
    hipStream_t st = (hipStream_t)stream;
    if (op->st.star_r > 0)   // one pass: out = L in + in * v
        return gpaw_check(launch_stencil(op->st, dtype, ST_HPSI, in, out, v, op->g, ph, (int)nin,
                                         0.0, nullptr, nullptr, true, st),
                          "operator apply_potential");
    // other stencils: the stencil, then the potential (same arithmetic)
    hipError_t e = launch_stencil(op->st, dtype, ST_APPLY, in, out, nullptr, op->g, ph, (int)nin,
                                  0.0, nullptr, nullptr, true, st);
    if (e == hipSuccess)
        e = launch_add_potential(out, in, v, nin,
                                 (long long)op->g.n[0] * op->g.n[1] * op->g.n[2], dtype, st);
    return gpaw_check(e, "operator apply_potential");
}

// Jacobi sweeps on f in place: ping-pong through `scratch`, ending in f.
int gpaw_op_jacobi(gpaw_operator* op, double* f, const double* s, int nrelax, double w,
                   double* scratch, hipStream_t st)
{
    Phases ph;
    gpaw_phases(nullptr, ph);
    const long long n = (long long)op->g.n[0] * op->g.n[1] * op->g.n[2];
    double* src = f;
    double* dst = scratch;
    if (nrelax % 2) {  // odd count: start from a copy so that the last sweep lands in f
        hipError_t e = launch_copy(scratch, f, DT_F64, n, st);
        if (e != hipSuccess) return gpaw_check(e, "relax copy");
        src = scratch;
        dst = f;
    }
    for (int it = 0; it < nrelax; it++) {
        hipError_t e = launch_stencil(op->st, DT_F64, ST_JACOBI, src, dst, s, op->g, ph, 1, w,
                                      nullptr, nullptr, true, st);
        if (e != hipSuccess) return gpaw_check(e, "relax sweep");
        double* t = src;
        src = dst;
        dst = t;
    }
    return GPAW_OK;
}

extern "C" int gpaw_operator_relax(gpaw_operator* op, int method, double* f, const double* s,
                                   int nrelax, double w, void* stream)
{
    if (!op) return gpaw_fail(GPAW_EINVAL, "operator: NULL handle");
    if (!op->real) return gpaw_fail(GPAW_EINVAL, "relax: operator is complex");
    if (!f || !s) return gpaw_fail(GPAW_EINVAL, "relax: NULL array");
    if (nrelax <= 0) return GPAW_OK;
    if (method != 1 && method != 2) return gpaw_fail(GPAW_EINVAL, "relax: unknown method %d", method);
    int rc = gpaw_op_ensure_scratch(op);
    if (rc) return rc;
    if (method == 1) {
        if (!op->st.d_coefs) {   // star stencils upload their table on first use
            hipError_t e = hipMalloc(&op->st.d_coefs, sizeof(double) * op->st.nc);
            if (e == hipSuccess) e = hipMalloc(&op->st.d_off, sizeof(int) * 3 * op->st.nc);
            if (e == hipSuccess)
                e = hipMemcpy(op->st.d_coefs, op->st.coefs, sizeof(double) * op->st.nc,
                              hipMemcpyHostToDevice);
            if (e == hipSuccess)
                e = hipMemcpy(op->st.d_off, op->st.off, sizeof(int) * 3 * op->st.nc,
                              hipMemcpyHostToDevice);
            if (e != hipSuccess) return gpaw_check(e, "relax: stencil upload");
        }
        return gpaw_op_gauss_seidel(op, f, s, nrelax, op->scratch, (hipStream_t)stream);
    }
    return gpaw_op_jacobi(op, f, s, nrelax, w, op->scratch, (hipStream_t)stream);
}

extern "C" double gpaw_operator_get_diagonal_element(const gpaw_operator* op)
{
    return op ? op->diag : 0.0;
}

int gpaw_op_residual(gpaw_operator* op, const double* phi, const double* rho, double* r,
                     double* norm2, double scale, double* partial, hipStream_t st)
{
    Phases ph;
    gpaw_phases(nullptr, ph);
    int np = 0;
    hipError_t e = launch_stencil(op->st, DT_F64, ST_RESID, phi, r, rho, op->g, ph, 1, 0.0,
                                  norm2 ? partial : nullptr, &np, r != nullptr, st);
    if (e == hipSuccess && norm2) e = launch_reduce_partials(partial, np, scale, norm2, st);
    return gpaw_check(e, "residual");
}

extern "C" int gpaw_operator_residual(gpaw_operator* op, const double* phi, const double* rho,
                                      double* r, double* norm2, double scale, void* stream)
{
    if (!op) return gpaw_fail(GPAW_EINVAL, "operator: NULL handle");
    if (!op->real) return gpaw_fail(GPAW_EINVAL, "residual: operator is complex");
    if (!phi || !rho) return gpaw_fail(GPAW_EINVAL, "residual: NULL array");
    if (!r && !norm2) return GPAW_OK;
    if (norm2) {
        int rc = gpaw_op_ensure_partials(op);
        if (rc) return rc;
    }
    return gpaw_op_residual(op, phi, rho, r, norm2, scale, op->partial, (hipStream_t)stream);
}

extern "C" void gpaw_operator_free(gpaw_operator* op)
{
    if (!op) return;
    if (op->st.d_coefs) hipFree(op->st.d_coefs);
    if (op->st.d_off) hipFree(op->st.d_off);
    if (op->scratch) hipFree(op->scratch);
    if (op->partial) hipFree(op->partial);
    delete op;
}

// ----------------------------------------------------------------------------
extern "C" int gpaw_transformer_new(const long size_in[3], const long size_out[3], int k,
                                    const long* pad, const long* npad, const long* skip,
                                    const long* neighbors, int real, int interpolate,
                                    gpaw_transformer** out)
{
    if (!out || !size_in || !size_out || !pad || !npad || !skip || !neighbors)
        return gpaw_fail(GPAW_EINVAL, "transformer: NULL argument");
    *out = nullptr;
    if (k != 2 && k != 4 && k != 6 && k != 8)
        return gpaw_fail(GPAW_EINVAL, "transformer: k = %d (must be 2, 4, 6 or 8)", k);
    gpaw_transformer* tr = new gpaw_transformer();
    Transfer& t = tr->t;
    t.K = k;
    t.interpolate = interpolate ? 1 : 0;
    int rc = neighbors_to_periodic(neighbors, t.gin.per);
    if (rc) {
        delete tr;
        return rc;
    }
    for (int a = 0; a < 3; a++) {
        const long n = size_in[a], lo = pad[2 * a], hi = pad[2 * a + 1];
        if (n < 1 || lo < 0 || hi < 0) {
            delete tr;
            return gpaw_fail(GPAW_EINVAL, "transformer: axis %d size %ld pad (%ld,%ld)", a, n,
                             lo, hi);
        }
        const long m = n + lo + hi;
        long expect;
        if (interpolate) {
            if (skip[2 * a] < 0 || skip[2 * a] > 1 || skip[2 * a + 1] < 0 || skip[2 * a + 1] > 1) {
                delete tr;
                return gpaw_fail(GPAW_EINVAL, "transformer: skip must be 0 or 1");
            }
            expect = (m - (k - 1)) * 2 - skip[2 * a] + skip[2 * a + 1];
        } else {
            if ((m - (2 * k - 3)) % 2) {
                delete tr;
                return gpaw_fail(GPAW_EINVAL, "transformer: axis %d padded size %ld is not "
                                              "restrictable with k = %d", a, m, k);
            }
            expect = (m - (2 * k - 3)) / 2;
        }
        if (expect != size_out[a] || expect < 1) {
            delete tr;
            return gpaw_fail(GPAW_EINVAL,
                             "transformer: axis %d output size %ld, padding implies %ld", a,
                             size_out[a], expect);
        }
        if (t.gin.per[a]) {
            if (lo > n || hi > n) {
                delete tr;
                return gpaw_fail(GPAW_EUNSUPPORTED,
                                 "transformer: periodic axis %d of %ld points narrower than "
                                 "its padding (%ld,%ld)", a, n, lo, hi);
            }
            if (npad[2 * a] != hi || npad[2 * a + 1] != lo) {
                delete tr;
                return gpaw_fail(GPAW_EUNSUPPORTED,
                                 "transformer: periodic axis %d sends (%ld,%ld) but pads "
                                 "(%ld,%ld); only a self-wrapping domain is supported",
                                 a, npad[2 * a], npad[2 * a + 1], lo, hi);
            }
        }
        t.gin.n[a] = (int)n;
        t.nout[a] = (int)size_out[a];
        t.lo[a] = (int)lo;
        t.skip[a][0] = interpolate ? (int)skip[2 * a] : 0;
        t.skip[a][1] = interpolate ? (int)skip[2 * a + 1] : 0;
    }
    tr->real = real ? 1 : 0;
    *out = tr;
    return GPAW_OK;
}

extern "C" int gpaw_transformer_apply(gpaw_transformer* tr, const void* in, void* out, long nin,
                                      int dtype, const double* phases, void* stream)
{
    if (!tr) return gpaw_fail(GPAW_EINVAL, "transformer: NULL handle");
    int rc = gpaw_check_dtype(tr->real, dtype);
    if (rc) return rc;
    if (nin < 0 || nin > 65535) return gpaw_fail(GPAW_EINVAL, "transformer: nin = %ld", nin);
    if (nin == 0) return GPAW_OK;
    if (!in || !out) return gpaw_fail(GPAW_EINVAL, "transformer: NULL array");
    Phases ph;
    gpaw_phases(phases, ph);
    return gpaw_check(launch_transfer(tr->t, dtype, TR_STORE, in, out, nullptr, ph, (int)nin,
                                      (hipStream_t)stream),
                      "transformer apply");
}

extern "C" void gpaw_transformer_free(gpaw_transformer* tr) { delete tr; }

extern "C" int gpaw_sum(const double* a, long n, double scale, double* result, void* stream)
{
    if (!a || !result || n < 0) return gpaw_fail(GPAW_EINVAL, "sum: bad argument");
    double* partial = nullptr;
    hipError_t e = hipMallocAsync((void**)&partial, sizeof(double) * 1024, (hipStream_t)stream);
    if (e != hipSuccess) return gpaw_check(e, "sum: partials");
    int np = 0;
    e = launch_sum(a, n, partial, &np, (hipStream_t)stream);
    if (e == hipSuccess) e = launch_reduce_partials(partial, np, scale, result, (hipStream_t)stream);
    hipError_t e2 = hipFreeAsync(partial, (hipStream_t)stream);
    if (e == hipSuccess) e = e2;
    return gpaw_check(e, "sum");
}

extern "C" int gpaw_dots(const double* a, const double* b, long nb, long n, double scale,
                         double* result, void* stream)
{
    if (!a || !result || (!b && nb > 0) || n < 0 || nb < 0 || nb > 65535)
        return gpaw_fail(GPAW_EINVAL, "dots: bad argument (0 <= nb <= 65535)");
    if (nb == 0) return GPAW_OK;
    double* partial = nullptr;
    const size_t bytes = sizeof(double) * (size_t)nb * (size_t)dots_chunks(n);
    hipError_t e = hipMallocAsync((void**)&partial, bytes, (hipStream_t)stream);
    if (e != hipSuccess) return gpaw_check(e, "dots: partials");
    e = launch_dots(a, b, (int)nb, n, scale, partial, result, (hipStream_t)stream);
    hipError_t e2 = hipFreeAsync(partial, (hipStream_t)stream);
    if (e == hipSuccess) e = e2;
    return gpaw_check(e, "dots");
}

// Partials of gpaw_dots_noalloc: one file-scope device array (each GPU the
// library runs on holds its own copy until the process ends).  In host code
// its name only stands in for the symbol: the device address of this GPU's
// copy comes from hipGetSymbolAddress, once per device, and that address is
// what the kernels get.  Every call on a GPU shares that GPU's copy, so calls
// that overlap on one GPU race; each call writes every partial it reads.
#define GPAW_DOTS_SCRATCH (1 << 19)   // doubles (4 MiB)
static __device__ __align__(256) double scratch[GPAW_DOTS_SCRATCH];

static double* dots_scratch(int* rc)
{
    static std::mutex mu;
    static double* ptr[64] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        *rc = gpaw_check(e, "dots_noalloc: current device");
        return nullptr;
    }
    if (dev < 0 || dev >= 64) {
        *rc = gpaw_fail(GPAW_EUNSUPPORTED, "dots_noalloc: device %d beyond 63", dev);
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *rc = gpaw_check(e != hipSuccess ? e : hipErrorInvalidSymbol,
                             "dots_noalloc: scratch address");
            return nullptr;
        }
        ptr[dev] = (double*)p;
    }
    *rc = GPAW_OK;
    return ptr[dev];
}

extern "C" int gpaw_dots_noalloc(const double* a, const double* b, long nb, long n, double scale,
                                 double* result, void* stream)
{
    if (!a || !result || (!b && nb > 0) || n < 0 || nb < 0 || nb > 65535)
        return gpaw_fail(GPAW_EINVAL, "dots_noalloc: bad argument (0 <= nb <= 65535)");
    if (nb == 0) return GPAW_OK;
    const long need = nb * (long)dots_chunks(n);
    if (need > GPAW_DOTS_SCRATCH)
        return gpaw_fail(GPAW_EINVAL,
                         "dots_noalloc: %ld rows x %d chunks need %ld partials, the static "
                         "scratch holds %d (use gpaw_dots)",
                         nb, dots_chunks(n), need, GPAW_DOTS_SCRATCH);
    int rc = GPAW_OK;
    double* partial = dots_scratch(&rc);
    if (!partial) return rc;
    return gpaw_check(launch_dots(a, b, (int)nb, n, scale, partial, result, (hipStream_t)stream),
                      "dots_noalloc");
}

extern "C" int gpaw_axpby(double* y, const double* x, double a, double b, long n, void* stream)
{
    if (!y || (!x && b != 0.0) || n < 0) return gpaw_fail(GPAW_EINVAL, "axpby: bad argument");
    return gpaw_check(launch_axpby(y, x ? x : y, a, b, n, (hipStream_t)stream), "axpby");
}

extern "C" int gpaw_sub_scalar(double* y, const double* x, double b, long n, void* stream)
{
    if (!y || !x || n < 0) return gpaw_fail(GPAW_EINVAL, "sub_scalar: bad argument");
    return gpaw_check(launch_sub_scalar(y, x, b, n, (hipStream_t)stream), "sub_scalar");
}

extern "C" int gpaw_cast(void* y, int dty, const void* x, int dtx, long n, void* stream)
{
    if (!y || !x || n < 0) return gpaw_fail(GPAW_EINVAL, "cast: bad argument");
    return gpaw_check(launch_cast(y, dty, x, dtx, n, (hipStream_t)stream), "cast");
}

extern "C" int gpaw_copy(void* y, const void* x, long nbytes, void* stream)
{
    if (!y || !x || nbytes < 0) return gpaw_fail(GPAW_EINVAL, "copy: bad argument");
    if ((nbytes | (long)(uintptr_t)y | (long)(uintptr_t)x) & 15)
        return gpaw_fail(GPAW_EINVAL, "copy: sizes and addresses must be multiples of 16 bytes");
    if (nbytes == 0) return GPAW_OK;
    return gpaw_check(launch_copy_bytes(y, x, nbytes, (hipStream_t)stream), "copy");
}

extern "C" int gpaw_add_potential(void* out, const void* in, const double* v, long nin,
                                  long ng, int dtype, void* stream)
{
    if (!out || !in || !v || nin < 0 || ng < 0)
        return gpaw_fail(GPAW_EINVAL, "add_potential: bad argument");
    if (dtype != DT_F64 && dtype != DT_C128)
        return gpaw_fail(GPAW_EINVAL, "add_potential: float64 or complex128 bands");
    if (out == in) return gpaw_fail(GPAW_EINVAL, "add_potential: out must not alias in");
    return gpaw_check(launch_add_potential(out, in, v, nin, ng, dtype, (hipStream_t)stream),
                      "add_potential");
}

extern "C" int gpaw_scale(double* y, double a, long n, void* stream)
{
    if (!y || n < 0) return gpaw_fail(GPAW_EINVAL, "scale: bad argument");
    return gpaw_check(launch_scale(y, a, n, (hipStream_t)stream), "scale");
}

extern "C" int gpaw_autotune(int enable) { return autotune_set(enable); }

extern "C" long gpaw_autotune_report(char* buf, long len) { return autotune_report(buf, len); }

extern "C" int gpaw_tile_shape(int threads, int tz, int planes)
{
    if (tile_shape_set(threads, tz, planes))
        return gpaw_fail(GPAW_EINVAL, "tile shape: threads %d, columns %d, planes %d", threads, tz,
                         planes);
    return GPAW_OK;
}

extern "C" int gpaw_stencil_kernel(int kind)
{
    // header numbering: 0 default, 1 tile, 2 box, 3 queue, 4 generic
    static const int to_internal[5] = {-1, 2, 3, 1, 0};
    static const int to_header[5] = {4, 3, 1, 2, 0};   // internal 0..3 -> header
    if (kind < -1 || kind > 4) return -1;
    const int prev = stencil_kernel_set(kind < 0 ? -2 : to_internal[kind]);
    return prev < 0 ? 0 : to_header[prev];
}
