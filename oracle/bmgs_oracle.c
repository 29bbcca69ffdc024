/*
 * bmgs_oracle.c -- TEST INFRASTRUCTURE ONLY (the parity oracle).
 *
 * A plain-C, single-threaded CPU restatement of GPAW's real-space
 * finite-difference hot path:
 *
 *   - the halo plan and halo fill of c/bc.c (bc_init c/bc.c:17-135,
 *     bc_unpack1 c/bc.c:138-250; serial path, no MPI),
 *   - the stencil kernels of c/bmgs/ (fd.c:16-51, relax.c:7-73,
 *     restrict.c:18-63,137-154, interpolate.c:19-66,141-168),
 *   - the band loops of the Operator / Transformer objects
 *     (c/operators.c:49-91,109-157; c/transformers.c:54-123).
 *
 * Only the tests, the smoke check and the CPU-baseline leg of the benchmark
 * load this library; the product (gpaw_amd) never does.
 *
 * Every expression keeps the reference's operand order and grouping and the
 * file is built with -ffp-contract=off, so results are bit-identical to the
 * reference C built without FMA contraction.  Complex arrays are interleaved
 * (re, im) doubles; complex*complex products use (ac-bd, ad+bc), complex*real
 * products scale both parts, exactly as C99 complex arithmetic does for
 * finite operands.
 *
 * Arrays are C-contiguous [band][i0][i1][i2] with i2 fastest.
 */
#include <stdlib.h>
#include <string.h>

typedef long idx_t;

/* ------------------------------------------------------------------ */
/* Halo plan (restates bc_init, c/bc.c:17-135, for one domain).        */
/* ------------------------------------------------------------------ */
typedef struct {
    idx_t n[3];          /* interior extent                           */
    idx_t m[3];          /* padded extent  n + pad_lo + pad_hi         */
    idx_t lo[3];         /* pad_lo per axis                           */
    /* per axis i, side d: the box copied by the periodic self-wrap   */
    idx_t src[3][2][3];  /* source start (padded coordinates)         */
    idx_t dst[3][2][3];  /* destination start                         */
    idx_t ext[3][2][3];  /* box extent                                */
    int wrap[3][2];      /* 1 when the side copies (periodic, ds > 0)  */
} halo_plan;

static int plan_halo(halo_plan* P, const idx_t n[3], const idx_t pad[3][2],
                     const idx_t npad[3][2], const idx_t nb[3][2], int cfd)
{
    idx_t box_start[3], box_size[3];
    for (int a = 0; a < 3; a++) {
        P->n[a] = n[a];
        P->lo[a] = pad[a][0];
        P->m[a] = n[a] + pad[a][0] + pad[a][1];
        box_start[a] = pad[a][0];
        box_size[a] = n[a];
    }
    for (int a = 0; a < 3; a++) {
        for (int d = 0; d < 2; d++) {
            if (nb[a][d] != 0 && nb[a][d] != -1)
                return -1;      /* only "self" (periodic) or "none" exist */
            idx_t ds = npad[a][d];
            for (int b = 0; b < 3; b++) {
                P->src[a][d][b] = box_start[b];
                P->ext[a][d][b] = box_size[b];
            }
            /* slab of thickness ds taken from the opposite interior edge */
            P->src[a][d][a] = (d == 0) ? pad[a][0] : pad[a][0] + n[a] - ds;
            P->ext[a][d][a] = ds;
            P->wrap[a][d] = (nb[a][d] == 0 && ds > 0);
        }
        /* side d lands in the halo on side 1-d */
        for (int d = 0; d < 2; d++) {
            for (int b = 0; b < 3; b++)
                P->dst[a][d][b] = P->src[a][d][b];
            P->dst[a][d][a] = (d == 0) ? pad[a][0] + n[a] : 0;
        }
        if (!cfd) {             /* later axes copy whole padded slabs */
            box_start[a] = 0;
            box_size[a] = P->m[a];
        }
    }
    return 0;
}

/* buf <- padded copy of in, halos filled as bc_unpack1 does
   (zero everything, paste the interior, then per axis 0,1,2 copy the
   periodic slabs multiplied by phases[a][d]). */
static void fill_halo(const halo_plan* P, int nd, const double* in,
                      double* buf, const double* phases)
{
    const idx_t M0 = P->m[0], M1 = P->m[1], M2 = P->m[2];
    memset(buf, 0, sizeof(double) * nd * M0 * M1 * M2);
    for (idx_t i0 = 0; i0 < P->n[0]; i0++)
        for (idx_t i1 = 0; i1 < P->n[1]; i1++) {
            const double* s = in + nd * ((i0 * P->n[1] + i1) * P->n[2]);
            double* t = buf + nd * (((i0 + P->lo[0]) * M1 + i1 + P->lo[1]) * M2
                                    + P->lo[2]);
            memcpy(t, s, sizeof(double) * nd * P->n[2]);
        }
    for (int a = 0; a < 3; a++)
        for (int d = 0; d < 2; d++) {
            if (!P->wrap[a][d])
                continue;
            const idx_t* s0 = P->src[a][d];
            const idx_t* t0 = P->dst[a][d];
            const idx_t* e = P->ext[a][d];
            double pr = 1.0, pi = 0.0;
            if (nd == 2) {
                pr = phases[4 * a + 2 * d];
                pi = phases[4 * a + 2 * d + 1];
            }
            for (idx_t j0 = 0; j0 < e[0]; j0++)
                for (idx_t j1 = 0; j1 < e[1]; j1++)
                    for (idx_t j2 = 0; j2 < e[2]; j2++) {
                        idx_t ps = ((s0[0] + j0) * M1 + s0[1] + j1) * M2 + s0[2] + j2;
                        idx_t pt = ((t0[0] + j0) * M1 + t0[1] + j1) * M2 + t0[2] + j2;
                        if (nd == 1) {
                            buf[pt] = buf[ps];
                        } else {
                            double ar = buf[2 * ps], ai = buf[2 * ps + 1];
                            buf[2 * pt] = pr * ar - pi * ai;
                            buf[2 * pt + 1] = pr * ai + pi * ar;
                        }
                    }
        }
}

/* ------------------------------------------------------------------ */
/* Stencil apply (restates bmgs_fd / bmgs_fdz, c/bmgs/fd.c:16-51).     */
/* offsets are flat offsets into the padded buffer of extent n + 2r.   */
/* ------------------------------------------------------------------ */
static void stencil_apply(int nd, int nc, const double* coefs,
                          const idx_t* offsets, const idx_t n[3], int r,
                          const double* buf, double* out)
{
    const idx_t M1 = n[1] + 2 * r, M2 = n[2] + 2 * r;
    for (idx_t i0 = 0; i0 < n[0]; i0++)
        for (idx_t i1 = 0; i1 < n[1]; i1++)
            for (idx_t i2 = 0; i2 < n[2]; i2++) {
                idx_t p = ((i0 + r) * M1 + i1 + r) * M2 + i2 + r;
                idx_t q = (i0 * n[1] + i1) * n[2] + i2;
                for (int k = 0; k < nd; k++) {
                    double x = 0.0;
                    for (int c = 0; c < nc; c++)
                        x += buf[nd * (p + offsets[c]) + k] * coefs[c];
                    out[nd * q + k] = x;
                }
            }
}

/* ------------------------------------------------------------------ */
/* Smoother (restates bmgs_relax, c/bmgs/relax.c:7-73).                */
/* method 1: lexicographic Gauss-Seidel, updates buf and f in place.   */
/* method 2: weighted Jacobi, reads buf (old f), writes f.             */
/* ------------------------------------------------------------------ */
static void relax_sweep(int method, int nc, const double* coefs,
                        const idx_t* offsets, const idx_t n[3], int r,
                        double* buf, double* f, const double* s, double w)
{
    const idx_t M1 = n[1] + 2 * r, M2 = n[2] + 2 * r;
    const double inv0 = 1.0 / coefs[0];
    for (idx_t i0 = 0; i0 < n[0]; i0++)
        for (idx_t i1 = 0; i1 < n[1]; i1++)
            for (idx_t i2 = 0; i2 < n[2]; i2++) {
                idx_t p = ((i0 + r) * M1 + i1 + r) * M2 + i2 + r;
                idx_t q = (i0 * n[1] + i1) * n[2] + i2;
                double x = 0.0;
                for (int c = 1; c < nc; c++)
                    x += buf[p + offsets[c]] * coefs[c];
                if (method == 1) {
                    x = (s[q] - x) * inv0;
                    f[q] = x;
                    buf[p] = x;
                } else {
                    f[q] = (1.0 - w) * f[q] + w * (s[q] - x) / coefs[0];
                }
            }
}

/* ------------------------------------------------------------------ */
/* 2x restriction (restates RST1DW / bmgs_restrict,                     */
/* c/bmgs/restrict.c:18-63,137-154): three 1-D passes z, y, x; each    */
/* pass reads rows of length 2*nout + 2K - 3 and writes transposed.     */
/* ------------------------------------------------------------------ */
static double rst_tap(int K, const double* c, int nd)
{
    /* c points at the centre sample; samples are nd doubles apart */
#define A(o) c[(o) * nd]
    switch (K) {
    case 2:
        return 0.5 * (A(0) + 0.5 * (A(1) + A(-1)));
    case 4:
        return 0.5 * (A(0) + 0.5625 * (A(1) + A(-1))
                      + -0.0625 * (A(3) + A(-3)));
    case 6:
        return 0.5 * (A(0) + 0.58593750 * (A(1) + A(-1))
                      + -0.09765625 * (A(3) + A(-3))
                      + 0.01171875 * (A(5) + A(-5)));
    default:
        return 0.5 * (A(0) + 0.59814453125 * (A(1) + A(-1))
                      + -0.11962890625 * (A(3) + A(-3))
                      + 0.02392578125 * (A(5) + A(-5))
                      + -0.00244140625 * (A(7) + A(-7)));
    }
#undef A
}

static void restrict_pass(int K, int nd, const double* a, idx_t nout,
                          idx_t rows, double* b)
{
    const idx_t len = 2 * nout + 2 * K - 3;
    for (idx_t j = 0; j < rows; j++)
        for (idx_t i = 0; i < nout; i++)
            for (int k = 0; k < nd; k++)
                b[nd * (i * rows + j) + k] =
                    rst_tap(K, a + nd * (j * len + K - 1 + 2 * i) + k, nd);
}

static void restrict3d(int K, int nd, double* buf, const idx_t m[3],
                       double* out, double* w)
{
    const idx_t e = 2 * K - 3;
    const idx_t o2 = (m[2] - e) / 2, o1 = (m[1] - e) / 2, o0 = (m[0] - e) / 2;
    restrict_pass(K, nd, buf, o2, m[0] * m[1], w);
    restrict_pass(K, nd, w, o1, m[0] * o2, buf);
    restrict_pass(K, nd, buf, o0, o1 * o2, out);
}

/* ------------------------------------------------------------------ */
/* 2x interpolation (restates IP1DW / bmgs_interpolate,                 */
/* c/bmgs/interpolate.c:19-66,141-168): three 1-D passes z, y, x; even */
/* outputs copy, odd outputs are the symmetric Lagrange midpoint; skip */
/* drops the first / last half-point on non-periodic axes.             */
/* ------------------------------------------------------------------ */
static double ip_tap(int K, const double* c, int nd)
{
#define A(o) c[(o) * nd]
    switch (K) {
    case 2:
        return 0.5 * (A(0) + A(1));
    case 4:
        return (0.5625 * (A(0) + A(1)) + -0.0625 * (A(-1) + A(2)));
    case 6:
        return (0.58593750 * (A(0) + A(1)) + -0.09765625 * (A(-1) + A(2))
                + 0.01171875 * (A(-2) + A(3)));
    default:
        return (0.59814453125 * (A(0) + A(1)) + -0.11962890625 * (A(-1) + A(2))
                + 0.02392578125 * (A(-2) + A(3))
                + -0.00244140625 * (A(-3) + A(4)));
    }
#undef A
}

static void interp_pass(int K, int nd, const double* a, idx_t n, idx_t rows,
                        double* b, const idx_t skip[2])
{
    const idx_t len = K - 1 - skip[1] + n;
    for (idx_t j = 0; j < rows; j++) {
        idx_t o = 0;                            /* next output position */
        for (idx_t i = 0; i < n; i++) {
            const double* c = a + nd * (j * len + K / 2 - 1 + i);
            if (!(i == 0 && skip[0])) {
                for (int k = 0; k < nd; k++)
                    b[nd * (o * rows + j) + k] = c[k];
                o++;
            }
            if (!(i == n - 1 && skip[1])) {
                for (int k = 0; k < nd; k++)
                    b[nd * (o * rows + j) + k] = ip_tap(K, c + k, nd);
                o++;
            }
        }
    }
}

static idx_t interp_out(int K, idx_t m, const idx_t skip[2])
{
    return (m - (K - 1)) * 2 - skip[0] + skip[1];
}

/* w1 holds the z-pass output (m0*m1*q2), w2 the y-pass output (m0*q1*q2) */
static void interpolate3d(int K, int nd, const idx_t skip[3][2],
                          const double* buf, const idx_t m[3], double* out,
                          double* w1, double* w2)
{
    const idx_t e = K - 1;
    const idx_t q2 = interp_out(K, m[2], skip[2]);
    const idx_t q1 = interp_out(K, m[1], skip[1]);
    interp_pass(K, nd, buf, m[2] - e + skip[2][1], m[0] * m[1], w1, skip[2]);
    interp_pass(K, nd, w1, m[1] - e + skip[1][1], m[0] * q2, w2, skip[1]);
    interp_pass(K, nd, w2, m[0] - e + skip[0][1], q1 * q2, out, skip[0]);
}

/* ------------------------------------------------------------------ */
/* Object-level drivers (stateless).                                   */
/* ------------------------------------------------------------------ */

/* Operator.apply: c/operators.c:316-405 with apply_worker :109-157.
   phases: complex128 [3][2] as 12 doubles, NULL for real data. */
int orc_operator_apply(const double* coefs, const long* offsets, int ncoefs,
                       const long size[3], int range, const long nb[3][2],
                       int real, int cfd, const double* in, double* out,
                       int nin, const double* phases)
{
    idx_t pad[3][2];
    for (int a = 0; a < 3; a++)
        pad[a][0] = pad[a][1] = range;
    halo_plan P;
    if (plan_halo(&P, size, pad, pad, nb, cfd))
        return -1;
    const int nd = real ? 1 : 2;
    const idx_t ng = nd * size[0] * size[1] * size[2];
    double* buf = malloc(sizeof(double) * nd * P.m[0] * P.m[1] * P.m[2]);
    if (!buf)
        return -2;
    for (int b = 0; b < nin; b++) {
        fill_halo(&P, nd, in + b * ng, buf, phases);
        stencil_apply(nd, ncoefs, coefs, offsets, size, range, buf, out + b * ng);
    }
    free(buf);
    return 0;
}

/* Operator.relax: c/operators.c:49-91 (real data, halo rebuilt before
   every sweep). */
int orc_operator_relax(const double* coefs, const long* offsets, int ncoefs,
                       const long size[3], int range, const long nb[3][2],
                       int cfd, int method, double* f, const double* s,
                       int nrelax, double w)
{
    idx_t pad[3][2];
    for (int a = 0; a < 3; a++)
        pad[a][0] = pad[a][1] = range;
    halo_plan P;
    if (plan_halo(&P, size, pad, pad, nb, cfd))
        return -1;
    double* buf = malloc(sizeof(double) * P.m[0] * P.m[1] * P.m[2]);
    if (!buf)
        return -2;
    for (int it = 0; it < nrelax; it++) {
        fill_halo(&P, 1, f, buf, NULL);
        relax_sweep(method, ncoefs, coefs, offsets, size, range, buf, f, s, w);
    }
    free(buf);
    return 0;
}

/* Transformer.apply: c/transformers.c:127-185 with transapply_worker
   :54-123 (halo plan always cfd = 0). */
int orc_transformer_apply(const long size_in[3], const long size_out[3],
                          int k, const long pad[3][2], const long npad[3][2],
                          const long skip[3][2], const long nb[3][2], int real,
                          int interpolate, const double* in, double* out,
                          int nin, const double* phases)
{
    halo_plan P;
    if (plan_halo(&P, size_in, pad, npad, nb, 0))
        return -1;
    const int nd = real ? 1 : 2;
    const idx_t ng = nd * size_in[0] * size_in[1] * size_in[2];
    const idx_t ngout = nd * size_out[0] * size_out[1] * size_out[2];
    const idx_t ng2 = nd * P.m[0] * P.m[1] * P.m[2];
    double* buf = malloc(sizeof(double) * ng2);
    /* scratch for the intermediate passes */
    idx_t nw1 = ng2, nw2 = 0;
    if (interpolate) {
        idx_t q2 = interp_out(k, P.m[2], skip[2]);
        idx_t q1 = interp_out(k, P.m[1], skip[1]);
        nw1 = nd * P.m[0] * P.m[1] * q2;
        nw2 = nd * P.m[0] * q1 * q2;
    }
    double* w = malloc(sizeof(double) * (nw1 + nw2));
    if (!buf || !w) {
        free(buf);
        free(w);
        return -2;
    }
    for (int b = 0; b < nin; b++) {
        fill_halo(&P, nd, in + b * ng, buf, phases);
        if (interpolate)
            interpolate3d(k, nd, skip, buf, P.m, out + b * ngout, w, w + nw1);
        else
            restrict3d(k, nd, buf, P.m, out + b * ngout, w);
    }
    free(buf);
    free(w);
    return 0;
}
