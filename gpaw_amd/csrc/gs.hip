// Gauss-Seidel relaxation (bmgs_relax method 1, c/bmgs/relax.c:11-45) on
// gfx950, bit-identical to the reference's lexicographic sweep.
//
// The reference updates points in lexicographic (i0, i1, i2) order in
// place; a neighbour inside the grid is read updated if it comes earlier
// in that order, and halo values are those of the padded buffer filled at
// the start of the sweep (c/operators.c:75-86).  Ordering the points by
// hyperplanes t = w0 i0 + w1 i1 + i2 with every lexicographically earlier
// neighbour on a lower hyperplane and every later one on a higher
// hyperplane reproduces that exactly while all points of a hyperplane are
// updated in parallel:
//   star stencils (offsets along one axis only): w = (1, 1, 1);
//   any stencil of range 1:                      w = (4, 2, 1).
// One launch per hyperplane; halo reads come from a copy of f taken at the
// start of each sweep.
#include "capi_internal.h"

namespace gpaw {

__global__ __launch_bounds__(256) void k_gs_plane(double* __restrict__ f,
                                                  const double* __restrict__ snap,
                                                  const double* __restrict__ src, Geom g,
                                                  const double* __restrict__ coefs,
                                                  const int* __restrict__ off, int nc, int w0,
                                                  int w1, int t, double inv0)
{
    const int idx = blockIdx.x * 256 + threadIdx.x;
    const int n1 = g.n[1], n2 = g.n[2];
    if (idx >= n1 * n2) return;
    const int i1 = idx / n2, i2 = idx % n2;
    const int r = t - w1 * i1 - i2;
    if (r < 0 || r % w0) return;
    const int i0 = r / w0;
    if (i0 >= g.n[0]) return;
    double x = 0.0;
    for (int c = 1; c < nc; c++) {
        int j0 = i0 + off[3 * c], j1 = i1 + off[3 * c + 1], j2 = i2 + off[3 * c + 2];
        const bool inside = j0 >= 0 && j0 < g.n[0] && j1 >= 0 && j1 < n1 && j2 >= 0 && j2 < n2;
        double v;
        if (inside) {
            v = f[((i64)j0 * n1 + j1) * n2 + j2];
        } else {
            const int a = wrap_index(j0, g.n[0], g.per[0]);
            const int b = wrap_index(j1, n1, g.per[1]);
            const int d = wrap_index(j2, n2, g.per[2]);
            v = (a == 3 || b == 3 || d == 3) ? 0.0 : snap[((i64)j0 * n1 + j1) * n2 + j2];
        }
        x += v * coefs[c];
    }
    const i64 p = ((i64)i0 * n1 + i1) * n2 + i2;
    f[p] = (src[p] - x) * inv0;
}

}  // namespace gpaw

using namespace gpaw;

// hyperplane weights for a stencil, 0 if unsupported
static int gs_weights(const Stencil& s, int& w0, int& w1)
{
    bool star = true;
    for (int c = 0; c < s.nc; c++) {
        int nz = (s.off[c][0] != 0) + (s.off[c][1] != 0) + (s.off[c][2] != 0);
        if (nz > 1) star = false;
    }
    if (star) {
        w0 = w1 = 1;
        return 1;
    }
    if (s.range <= 1) {
        w0 = 4;
        w1 = 2;
        return 1;
    }
    return 0;
}

int gpaw_op_gauss_seidel(gpaw_operator* op, double* f, const double* s, int nrelax,
                         double* snap, hipStream_t st)
{
    int w0, w1;
    if (!gs_weights(op->st, w0, w1))
        return gpaw_fail(GPAW_EUNSUPPORTED,
                         "relax: Gauss-Seidel needs a star stencil or a range-1 stencil");
    const Geom& g = op->g;
    const long long n = (long long)g.n[0] * g.n[1] * g.n[2];
    const int tmax = w0 * (g.n[0] - 1) + w1 * (g.n[1] - 1) + (g.n[2] - 1);
    const int nb = (g.n[1] * g.n[2] + 255) / 256;
    const double inv0 = 1.0 / op->st.coefs[0];
    const bool periodic = g.per[0] || g.per[1] || g.per[2];
    for (int it = 0; it < nrelax; it++) {
        if (periodic) {
            hipError_t e = launch_copy(snap, f, DT_F64, n, st);
            if (e != hipSuccess) return gpaw_check(e, "relax: snapshot");
        }
        for (int t = 0; t <= tmax; t++)
            launch_kernel(k_gs_plane, dim3(nb), dim3(256), 0, st, f, snap, s, g,
                               op->st.d_coefs, op->st.d_off, op->st.nc, w0, w1, t, inv0);
        hipError_t e = launch_error();
        if (e != hipSuccess) return gpaw_check(e, "relax: Gauss-Seidel sweep");
    }
    return GPAW_OK;
}
