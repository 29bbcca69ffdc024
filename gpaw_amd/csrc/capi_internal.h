// Object layouts and helpers shared by the C-ABI translation units.
#pragma once
#include "../../include/gpaw_amd.h"
#include "kernels.h"

struct gpaw_operator {
    gpaw::Stencil st;
    gpaw::Geom g;
    int real = 1;
    int cfd = 1;
    int device = 0;
    double diag = 0.0;
    double* scratch = nullptr;   // Jacobi ping-pong buffer (n doubles)
    double* partial = nullptr;   // per-block norm partials
    int partial_n = 0;
};

struct gpaw_transformer {
    gpaw::Transfer t;
    int real = 1;
};

int gpaw_fail(int code, const char* fmt, ...);
int gpaw_check(hipError_t e, const char* what);
void gpaw_phases(const double* p, gpaw::Phases& ph);
int gpaw_check_dtype(int real, int dtype);
int gpaw_op_ensure_partials(gpaw_operator* op);
int gpaw_op_ensure_scratch(gpaw_operator* op);
int gpaw_op_jacobi(gpaw_operator* op, double* f, const double* s, int nrelax, double w,
                   double* scratch, hipStream_t st);
int gpaw_op_gauss_seidel(gpaw_operator* op, double* f, const double* s, int nrelax,
                         double* snap, hipStream_t st);
int gpaw_op_residual(gpaw_operator* op, const double* phi, const double* rho, double* r,
                     double* norm2, double scale, double* partial, hipStream_t st);
