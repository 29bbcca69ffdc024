/*
 * gpaw_amd.h -- C ABI of the MI355X (gfx950) finite-difference / multigrid
 * library (libgpaw_amd.so).
 *
 * The entry points replace the reference's compiled Python-extension
 * objects one for one; the citation on each names the reference interface
 * it stands in for (paths relative to the reference tree).  All array
 * arguments are DEVICE pointers (e.g. torch tensor data_ptr()), C-contiguous
 * [band][i0][i1][i2].  `stream` is a hipStream_t (NULL = default stream);
 * work is stream-ordered and the calls return without synchronising.
 * Every function returns 0 on success or a negative GPAW_E* code, and
 * gpaw_last_error() describes the last failure of the calling thread.
 *
 * Halo semantics are those of the reference's single-domain bc plan
 * (c/bc.c:17-250): neighbours[a][d] == 0 means "periodic self-wrap"
 * (the domain is its own neighbour), -1 means "zero boundary"; other ranks
 * (MPI domain decomposition) are not supported on one GPU.
 */
#ifndef GPAW_AMD_H
#define GPAW_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* storage types of grid functions */
#define GPAW_F64 0   /* float64                       */
#define GPAW_C128 1  /* complex128 (interleaved re,im) */
#define GPAW_F32 2   /* float32                       */
#define GPAW_C64 3   /* complex64                     */

#define GPAW_OK 0
#define GPAW_EINVAL (-1)       /* bad argument / inconsistent sizes      */
#define GPAW_EUNSUPPORTED (-2) /* valid in the reference, not on a GPU   */
#define GPAW_EDEVICE (-3)      /* HIP runtime error                      */
#define GPAW_ENOMEM (-4)       /* device allocation failed               */
#define GPAW_ENOCONV (-5)      /* Poisson solver did not converge        */

typedef struct gpaw_operator gpaw_operator;
typedef struct gpaw_transformer gpaw_transformer;
typedef struct gpaw_poisson gpaw_poisson;

const char* gpaw_last_error(void);

/* ---------------------------------------------------------------- Operator
 * Replaces _gpaw.Operator(coefs, offsets, size, range, neighbors, real,
 * comm, cfd): NewOperatorObject, c/operators.c:465-498 (Python caller
 * gpaw/fd_operators.py:57-74).  `offsets` are the reference's flat offsets
 * into the padded buffer of extent size + 2*range
 * (gpaw/fd_operators.py:39-42); comm must be None on one GPU and is not
 * passed.  neighbors is long[3][2]. */
int gpaw_operator_new(const double* coefs, const long* offsets, int ncoefs,
                      const long size[3], int range, const long* neighbors,
                      int real, int cfd, gpaw_operator** op);

/* Operator.apply(in, out, phases): Operator_apply, c/operators.c:316-405.
 * nin bands of `dtype`; phases = complex128 phase_cd[3][2] as 12 doubles
 * in HOST memory (NULL: all phases 1). */
int gpaw_operator_apply(gpaw_operator* op, const void* in, void* out,
                        long nin, int dtype, const double* phases,
                        void* stream);

/* out = L in + in * v for nin bands: the kinetic stencil and the local
 * potential v (float64, shared by the bands) in one pass.  Replaces the pair
 * kin.apply(psit_xG, Htpsit_xG, phase_cd) + apply_local_potential
 * (gpaw/wavefunctions/fd.py:51-52, gpaw/hamiltonian.py:393-398), with the
 * same rounding (the stencil sum, then + psi * v).  dtype GPAW_F64 or
 * GPAW_C128; in != out. */
int gpaw_operator_apply_potential(gpaw_operator* op, const void* in, void* out,
                                  long nin, int dtype, const double* phases,
                                  const double* v, void* stream);

/* Operator.relax(method, f, s, nrelax, w): Operator_relax,
 * c/operators.c:49-91 -> bmgs_relax, c/bmgs/relax.c:7-73.  float64 only;
 * f is updated in place.  method 2 = weighted Jacobi, 1 = lexicographic
 * Gauss-Seidel (star stencils and stencils of range 1; w is ignored, as in
 * the reference). */
int gpaw_operator_relax(gpaw_operator* op, int method, double* f,
                        const double* s, int nrelax, double w, void* stream);

/* Operator.get_diagonal_element(): c/operators.c:408-421. */
double gpaw_operator_get_diagonal_element(const gpaw_operator* op);

/* Fused residual r = L phi - rho of the multigrid loop
 * (gpaw/poisson.py:257-258,271-275: operator.apply then numpy r -= rho).
 * r may be NULL (not stored).  If norm2 is not NULL it receives, in DEVICE
 * memory, scale * sum(r^2) (the reference's dv * dot(r, r)). */
int gpaw_operator_residual(gpaw_operator* op, const double* phi,
                           const double* rho, double* r, double* norm2,
                           double scale, void* stream);

void gpaw_operator_free(gpaw_operator* op);

/* ------------------------------------------------------------ Transformer
 * Replaces _gpaw.Transformer(size_in, size_out, k, paddings, npaddings,
 * skip, neighbors, real, comm, interpolate): NewTransformerObject,
 * c/transformers.c:222-265 (Python caller gpaw/transformers.py:64-79).
 * paddings, npaddings, skip, neighbors are long[3][2]. */
int gpaw_transformer_new(const long size_in[3], const long size_out[3],
                         int k, const long* paddings, const long* npaddings,
                         const long* skip, const long* neighbors, int real,
                         int interpolate, gpaw_transformer** tr);

/* Transformer.apply(in, out, phases): Transformer_apply,
 * c/transformers.c:127-185. */
int gpaw_transformer_apply(gpaw_transformer* tr, const void* in, void* out,
                           long nin, int dtype, const double* phases,
                           void* stream);

void gpaw_transformer_free(gpaw_transformer* tr);

/* ------------------------------------------------------ multigrid Poisson
 * The V-cycle loop of PoissonSolver.solve_neutral / iterate2
 * (gpaw/poisson.py:193-218,245-276), run on the device.  The reference
 * drives it from Python through the Operator / Transformer objects above;
 * here one plan owns the level arrays and replays the V-cycle.
 * ops[0..levels], restrictors/interpolators[0..levels-1] are borrowed and
 * must outlive the plan; presmooths/postsmooths/weights have levels+1
 * entries (gpaw/poisson.py:64-89,111-113). */
int gpaw_poisson_new(int levels, gpaw_operator* const* ops,
                     gpaw_transformer* const* restrictors,
                     gpaw_transformer* const* interpolators,
                     const int* presmooths, const int* postsmooths,
                     const double* weights, int relax_method, double dv,
                     gpaw_poisson** ps);

/* solve_neutral(phi, rho, eps): V-cycles until dv*sum(r^2) <= eps or
 * maxiter cycles (niter counts as the reference does).  phi is updated in
 * place (device), rho is read.  Returns GPAW_ENOCONV when maxiter is
 * reached, with *niter and *error set. */
int gpaw_poisson_solve_neutral(gpaw_poisson* ps, double* phi,
                               const double* rho, double eps, int maxiter,
                               int* niter, double* error, void* stream);

/* The start of PoissonSolver.solve (gpaw/poisson.py:124-138): *charge =
 * dv * sum(rho) (gd.integrate); if |charge| <= maxcharge, rho minus its
 * mean (charge / dv / npts) goes into the plan's level-0 density and
 * *neutral = 1 (a following gpaw_poisson_solve_mixed may then pass
 * rho = NULL); else *neutral = 0 and nothing else is done (the caller's
 * charged branches).  One host round trip for the charge; the same kernels
 * and bits as the reference's steps from Python. */
int gpaw_poisson_neutralize(gpaw_poisson* ps, const double* rho, double maxcharge,
                            double* charge, int* neutral, void* stream);

/* gpaw_poisson_neutralize, then for a neutral density solve_neutral as
 * above (*neutral = 1); a charged one returns with *neutral = 0. */
int gpaw_poisson_solve(gpaw_poisson* ps, double* phi, const double* rho, double eps,
                       int maxiter, double maxcharge, double* charge, int* neutral,
                       int* niter, double* error, void* stream);

/* Exactly ncycles V-cycles; errors[i] (host) gets dv*sum(r^2) after
 * cycle i.  Used for fixed-cycle comparisons. */
int gpaw_poisson_vcycles(gpaw_poisson* ps, double* phi, const double* rho,
                         int ncycles, double* errors, void* stream);

/* One inner V-cycle of the mixed solver applied to r, from z = 0: the
 * multigrid preconditioner z ~ L^-1 r of its flexible CG, for callers that
 * run their own Krylov loop.  No reference counterpart (gpaw/poisson.py
 * iterates V-cycles on phi only).  The reference hierarchy and weights
 * (gpaw/poisson.py:64-113) with the mixed solver's symmetric level-0
 * smoothing (gpaw_poisson_mixed_weights); inner_precision 0: every level's
 * sweeps, residuals and transfers in FP32 (the smallest levels in FP64),
 * 1: all in FP64.  r, z: n0*n1*n2 doubles on the device, distinct; the
 * plan must use Jacobi smoothing.  Asynchronous on `stream`. */
int gpaw_poisson_precondition(gpaw_poisson* ps, const double* r, double* z,
                              int inner_precision, void* stream);

/* The smoothing of level `level` in the mixed solver's inner V-cycle (no
 * reference counterpart): *npre pre- and *npost post-smoothing Jacobi
 * sweeps and the weight of each, in the order they run (pre / post: room
 * for 16 each, or null to query the counts).  Level 0 by default: 2 + 2
 * sweeps whose weights make the product of the sweeps' factors
 * prod (1 - w_i D^-1 A) the degree-2 Chebyshev polynomial smallest on
 * [lmax/7, lmax] (lmax the largest eigenvalue of D^-1 A), the
 * post-smoothing weights in reverse order.  Level 1, when its fused
 * smoothing launches do two sweeps each (a level 1 of at least 73 points
 * along each axis, e.g. 79^3 under 159^3): 2 + 2 sweeps on [lmax/6, lmax] the same way.  The other levels
 * keep the reference's sweeps and weight. */
int gpaw_poisson_mixed_weights(const gpaw_poisson* ps, int level, double* pre, int* npre,
                               double* post, int* npost);

/* Mixed-precision solve of the same problem (no reference counterpart: the
 * FP32 / FP64 split of the configuration "mixed-precision multigrid
 * Poisson").  FP64 outer loop (residual L phi - rho, correction
 * accumulation, dot products, convergence) around inner V-cycles
 * (smoothing, residuals, restriction, interpolation; the fused coarsest
 * levels in FP64).  mode bit 0: 0 = iterative refinement with `inner`
 * V-cycles per step, 1 = flexible preconditioned CG with one V-cycle per
 * iteration; mode bit 1: 0 = inner V-cycle in FP32, 1 = the same inner
 * V-cycle in FP64 (for attributing the speed-up); mode bit 2 (CG only):
 * the first chunk of iterations is as long as the previous converged solve
 * on this plan to the same rel_tol took, ending with the true residual, so that a repeated
 * solve of a similar density reads a norm back once (the iteration count,
 * and so the last bits of phi, then depend on that history; without the
 * bit they depend only on the inputs).  Stops when
 * ||L phi - rho||_2 <= rel_tol * ||rho||_2 (rho as given, already
 * neutral; NULL: the density gpaw_poisson_neutralize left in the plan),
 * checked on the true residual; *relres receives the final
 * ratio.  A non-finite residual norm (breakdown) returns GPAW_EDEVICE and
 * leaves phi unchanged. */
int gpaw_poisson_solve_mixed(gpaw_poisson* ps, double* phi,
                             const double* rho, double rel_tol, int maxiter,
                             int mode, int inner, int* niter, double* relres,
                             void* stream);

void gpaw_poisson_free(gpaw_poisson* ps);

/* ---------------------------------------------------------- preconditioner
 * Replaces the Python class Preconditioner (gpaw/preconditioner.py:13-70),
 * whose __call__ drives _gpaw.Operator / _gpaw.Transformer objects and
 * numpy updates band-block by band-block; here the whole sequence runs on
 * the device for all nin bands in ten launches.  kin0 is the fine-grid
 * kinetic operator, kin1 / kin2 the nn=1 Laplacians (scale -0.5) of the two
 * coarser grids, restrictor0/1 and interpolator2/1 the nn=1 transformers
 * 0->1, 1->2, 2->1, 1->0 (gpaw/preconditioner.py:14-31).  All are borrowed
 * and must outlive the preconditioner.  step = 0.66666666 / kin0's
 * diagonal element (gpaw/preconditioner.py:24). */
typedef struct gpaw_preconditioner gpaw_preconditioner;

int gpaw_preconditioner_new(gpaw_operator* kin0, gpaw_operator* kin1,
                            gpaw_operator* kin2, gpaw_transformer* restrictor0,
                            gpaw_transformer* restrictor1,
                            gpaw_transformer* interpolator2,
                            gpaw_transformer* interpolator1,
                            gpaw_preconditioner** pc);

double gpaw_preconditioner_step(const gpaw_preconditioner* pc);

/* Device workspace (bytes) that gpaw_preconditioner_apply needs for nin
 * bands of dtype (GPAW_F64 or GPAW_C128): one fine and six coarse grid
 * functions per band (the reference's scratch0/1/2 arrays). */
int gpaw_preconditioner_workspace(const gpaw_preconditioner* pc, long nin,
                                  int dtype, long* bytes);

/* Preconditioner.__call__(residuals, kpt): out = the preconditioned
 * residuals for nin bands with the k-point's phase_cd (12 doubles in host
 * memory, NULL = Gamma).  out must not alias residuals; work is a device
 * buffer of at least gpaw_preconditioner_workspace bytes, 256-byte
 * aligned, owned by the caller (no allocation happens in this call). */
int gpaw_preconditioner_apply(gpaw_preconditioner* pc, const void* residuals,
                              void* out, long nin, int dtype,
                              const double* phases, void* work,
                              long work_bytes, void* stream);

void gpaw_preconditioner_free(gpaw_preconditioner* pc);

/* ----------------------------------------------------------- batched apply
 * No reference counterpart: the reference applies an operator to the bands
 * of one grid and one k-point per call (Operator_apply, c/operators.c:
 * 316-405).  Here several such calls -- bands of several k-points, grids of
 * uneven size, different coefficients -- run as one persistent launch per
 * stencil range.  Group i applies ops[i] to nins[i] bands at ins[i] into
 * outs[i] with the phase_cd of phases[12*i .. 12*i+11] (host memory; NULL
 * = all phases 1).  Results are bit-identical to gpaw_operator_apply of
 * each group.  All groups run concurrently, so no output may overlap any
 * group's input or another output (GPAW_EINVAL otherwise); inputs may be
 * shared.  A batch object holds the descriptor buffers for up to
 * max_groups groups; it may be reused, one call at a time. */
typedef struct gpaw_batch gpaw_batch;

int gpaw_batch_new(int max_groups, gpaw_batch** batch);

int gpaw_batch_apply(gpaw_batch* batch, int ngroups, gpaw_operator* const* ops,
                     const void* const* ins, void* const* outs,
                     const long* nins, const double* phases, int dtype,
                     void* stream);

void gpaw_batch_free(gpaw_batch* batch);

/* --------------------------------------------------------------- autotune
 * No reference counterpart (the reference's stencil loops have no launch
 * shape).  With the tuner on, the first apply / relax / residual launch of
 * each (dtype, mode, stencil range, grid, band count) times the candidate
 * launch shapes of the star-stencil kernels on that very launch and keeps
 * the fastest; every shape gives bit-identical results.  A Poisson plan
 * runs one V-cycle on its own buffers before recording its graphs, so that
 * they use tuned shapes.  enable: 1 on, 0 off, -1 query.  Returns the
 * previous setting.  Off by default; GPAW_AMD_AUTOTUNE=1 also turns it on. */
int gpaw_autotune(int enable);

/* The tuner's choices as text, one line per combination; returns the full
 * length (buf may be NULL to query it). */
long gpaw_autotune_report(char* buf, long len);

/* No reference counterpart: the kernel family of star-stencil launches that
 * carry no partial sums, for parity tests of every family and timing
 * experiments.  kind: 0 the default choice (heuristic or tuner), 1 the
 * plane-sweeping tile kernel, 2 the one-shot box kernel, 3 the register-
 * queue column kernel, 4 the generic kernel; a family that cannot run a
 * launch (grid too small, stencil not a star) leaves it to the default;
 * -1 only queries.  Returns the previous setting.  Every family gives
 * bit-identical results. */
int gpaw_stencil_kernel(int kind);

/* Launch shape of the tile kernel, with gpaw_stencil_kernel(1), for launches
 * without partial sums (tests and timing experiments; no reference
 * counterpart): `threads` per block (256, 512 or 1024; 512 and 1024 for
 * complex data of range <= 3 only), `tz` columns (32 or 64) and `planes`
 * per sweep (>= 1; at least n0 sweeps the whole i0 axis in one run).  A
 * shape that cannot run a launch leaves the default; threads = 0 clears the
 * setting.  Returns GPAW_OK or GPAW_EINVAL.  Every shape gives bit-identical
 * results. */
int gpaw_tile_shape(int threads, int tz, int planes);

/* -------------------------------------------------------------- helpers */
/* Deterministic FP64 sum of n doubles times scale into *result (device):
 * gd.integrate / comm.sum of the reference (gpaw/grid_descriptor.py:197-208). */
int gpaw_sum(const double* a, long n, double scale, double* result,
             void* stream);

/* result[j] = scale * sum_g a[g] * b[j*n + g] for j < nb (0 <= nb <= 65535;
 * device pointers, FP64, deterministic order): one row of the pair-density
 * coupling integrals gd.integrate(rhot * phit) of the RPA omega matrix
 * (gpaw/lrtddft/omega_matrix.py:427, gpaw/grid_descriptor.py:197-208). */
int gpaw_dots(const double* a, const double* b, long nb, long n, double scale,
              double* result, void* stream);

/* gpaw_dots without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The partial sums live in a static device array of 2^19
 * doubles (4 MiB per GPU, held until the process ends), so nb times the
 * number of chunks of n (ceil(n / 1024), at most 256) must not exceed 2^19;
 * larger calls return GPAW_EINVAL.  Same kernels, same results as
 * gpaw_dots.  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams, other threads) race. */
int gpaw_dots_noalloc(const double* a, const double* b, long nb, long n,
                      double scale, double* result, void* stream);

/* y = a*y + b*x over n doubles (numpy's in-place updates around the
 * solver, e.g. phi += q*phi_gauss, gpaw/poisson.py:180,186). */
int gpaw_axpby(double* y, const double* x, double a, double b, long n,
               void* stream);

/* y = x - b over n doubles (rho - background, phi -= phi_ave:
 * gpaw/poisson.py:138,216). */
int gpaw_sub_scalar(double* y, const double* x, double b, long n,
                    void* stream);

/* Device copy of nbytes (numpy's b[:] = a between grid functions, e.g.
 * gpaw/poisson.py:197 rhos[0][:] = rho); sizes and addresses multiples of 16.
 * Also the kernel that calibrates the attainable HBM read + write rate. */
int gpaw_copy(void* y, const void* x, long nbytes, void* stream);

/* out[b][g] += in[b][g] * v[g] for nin bands of ng points: the local
 * potential v (float64, shared by the bands) applied after the kinetic
 * stencil (replaces gpaw/hamiltonian.py:393-398, Htpsit_G += psit_G * vt_G,
 * called from gpaw/wavefunctions/fd.py:50-52).  dtype GPAW_F64 or
 * GPAW_C128 (both parts scaled); out must not alias in. */
int gpaw_add_potential(void* out, const void* in, const double* v, long nin,
                       long ng, int dtype, void* stream);

/* y *= a over n doubles (numpy's in-place scaling, e.g. nt_sg *= x,
 * gpaw/density.py:234). */
int gpaw_scale(double* y, double a, long n, void* stream);

/* Precision conversion between GPAW_F64 <-> GPAW_F32 or GPAW_C128 <->
 * GPAW_C64 (n elements). */
int gpaw_cast(void* y, int dty, const void* x, int dtx, long n, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* GPAW_AMD_H */
