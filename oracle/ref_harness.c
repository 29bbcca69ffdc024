/*
 * ref_harness.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Links against the reference's own kernels, compiled from their sources in
 * place (c/bmgs/bmgs.c and c/bc.c of the reference tree, serial build, no
 * MPI), and drives them with the same band loops as the reference's Python
 * extension objects:
 *   Operator.apply   c/operators.c:316-405, apply_worker :109-157
 *   Operator.relax   c/operators.c:49-91
 *   Transformer.apply c/transformers.c:127-185, transapply_worker :54-123
 * The entry points have the same signatures as the oracle restatement in
 * bmgs_oracle.c (prefix ref_ instead of orc_), so tests can compare the two.
 * The result (oracle/_ref/libgpawref.so) is never part of the product.
 */
#include <stdlib.h>
#include "bc.h"

int ref_operator_apply(const double* coefs, const long* offsets, int ncoefs,
                       const long size[3], int range, const long nb[3][2],
                       int real, int cfd, const double* in, double* out,
                       int nin, const double* phases)
{
    bmgsstencil st = bmgs_stencil(ncoefs, coefs, offsets, range, size);
    const long pad[3][2] = {{range, range}, {range, range}, {range, range}};
    boundary_conditions* bc = bc_init(size, pad, pad, nb, MPI_COMM_NULL,
                                      real, cfd);
    const int ng = bc->ndouble * bc->size1[0] * bc->size1[1] * bc->size1[2];
    const int ng2 = bc->ndouble * bc->size2[0] * bc->size2[1] * bc->size2[2];
    double* buf = malloc(sizeof(double) * ng2);
    double* sbuf = malloc(sizeof(double) * (bc->maxsend + 1));
    double* rbuf = malloc(sizeof(double) * (bc->maxrecv + 1));
    MPI_Request rreq[2], sreq[2];
    const double_complex* ph = (const double_complex*)phases;
    for (int b = 0; b < nin; b++) {
        for (int i = 0; i < 3; i++) {
            bc_unpack1(bc, in + (long)b * ng, buf, i, rreq, sreq, rbuf, sbuf,
                       ph ? ph + 2 * i : NULL, 0, 1);
            bc_unpack2(bc, buf, i, rreq, sreq, rbuf, 1);
        }
        if (real)
            bmgs_fd(&st, buf, out + (long)b * ng);
        else
            bmgs_fdz(&st, (const double_complex*)buf,
                     (double_complex*)(out + (long)b * ng));
    }
    free(buf);
    free(sbuf);
    free(rbuf);
    free(bc);
    free(st.coefs);
    free(st.offsets);
    return 0;
}

int ref_operator_relax(const double* coefs, const long* offsets, int ncoefs,
                       const long size[3], int range, const long nb[3][2],
                       int cfd, int method, double* f, const double* s,
                       int nrelax, double w)
{
    bmgsstencil st = bmgs_stencil(ncoefs, coefs, offsets, range, size);
    const long pad[3][2] = {{range, range}, {range, range}, {range, range}};
    boundary_conditions* bc = bc_init(size, pad, pad, nb, MPI_COMM_NULL, 1,
                                      cfd);
    const int ng2 = bc->size2[0] * bc->size2[1] * bc->size2[2];
    double* buf = malloc(sizeof(double) * ng2);
    double* sbuf = malloc(sizeof(double) * (bc->maxsend + 1));
    double* rbuf = malloc(sizeof(double) * (bc->maxrecv + 1));
    MPI_Request rreq[2], sreq[2];
    for (int n = 0; n < nrelax; n++) {
        for (int i = 0; i < 3; i++) {
            bc_unpack1(bc, f, buf, i, rreq, sreq, rbuf, sbuf, NULL, 0, 1);
            bc_unpack2(bc, buf, i, rreq, sreq, rbuf, 1);
        }
        bmgs_relax(method, &st, buf, f, s, w);
    }
    free(buf);
    free(sbuf);
    free(rbuf);
    free(bc);
    free(st.coefs);
    free(st.offsets);
    return 0;
}

int ref_transformer_apply(const long size_in[3], const long size_out[3],
                          int k, const long pad[3][2], const long npad[3][2],
                          const long skip[3][2], const long nb[3][2], int real,
                          int interpolate, const double* in, double* out,
                          int nin, const double* phases)
{
    boundary_conditions* bc = bc_init(size_in, pad, npad, nb, MPI_COMM_NULL,
                                      real, 0);
    const int ng = bc->ndouble * bc->size1[0] * bc->size1[1] * bc->size1[2];
    const int ng2 = bc->ndouble * bc->size2[0] * bc->size2[1] * bc->size2[2];
    const int ngout = bc->ndouble * size_out[0] * size_out[1] * size_out[2];
    int iskip[3][2];
    for (int c = 0; c < 3; c++)
        for (int d = 0; d < 2; d++)
            iskip[c][d] = (int)skip[c][d];
    double* buf = malloc(sizeof(double) * ng2);
    /* same scratch sizing as transapply_worker (c/transformers.c:63-67) */
    int n2 = interpolate ? 16 * ng2 : ng2 / 2;
    double* buf2 = malloc(sizeof(double) * (n2 + 1));
    double* sbuf = malloc(sizeof(double) * (bc->maxsend + 1));
    double* rbuf = malloc(sizeof(double) * (bc->maxrecv + 1));
    MPI_Request rreq[2], sreq[2];
    const double_complex* ph = (const double_complex*)phases;
    for (int b = 0; b < nin; b++) {
        for (int i = 0; i < 3; i++) {
            bc_unpack1(bc, in + (long)b * ng, buf, i, rreq, sreq, rbuf, sbuf,
                       ph ? ph + 2 * i : NULL, 0, 1);
            bc_unpack2(bc, buf, i, rreq, sreq, rbuf, 1);
        }
        double* o = out + (long)b * ngout;
        if (real) {
            if (interpolate)
                bmgs_interpolate(k, iskip, buf, bc->size2, o, buf2);
            else
                bmgs_restrict(k, buf, bc->size2, o, buf2);
        } else {
            if (interpolate)
                bmgs_interpolatez(k, iskip, (double_complex*)buf, bc->size2,
                                  (double_complex*)o, (double_complex*)buf2);
            else
                bmgs_restrictz(k, (double_complex*)buf, bc->size2,
                               (double_complex*)o, (double_complex*)buf2);
        }
    }
    free(buf);
    free(buf2);
    free(sbuf);
    free(rbuf);
    free(bc);
    return 0;
}
