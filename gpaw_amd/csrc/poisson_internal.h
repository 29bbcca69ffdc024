// Layout of the device Poisson plan, shared by the FP64 V-cycle
// (poisson.hip) and the mixed-precision solver (mixed.hip).
#pragma once
#include <vector>

#define GPAW_CHUNK_MAX 32   // V-cycles per chunk

#include "capi_internal.h"
#include "coarse.h"

struct MixedState;

struct gpaw_poisson {
    int L = 0;                 // number of coarsenings
    int method = 2;
    double dv = 1.0;
    std::vector<gpaw_operator*> ops;
    std::vector<gpaw_transformer*> rst, itp;
    std::vector<int> pre, post;
    std::vector<double> w;
    std::vector<double*> phiA, phiB, rho, res;
    std::vector<long long> npts;
    double* d_err = nullptr;
    double* h_err = nullptr;
    double* err_dst = nullptr;  // where the level-0 norm of the next emitted cycle goes
    // when set: the level-0 norm partials of the next emitted cycle go here
    // and are summed later (one launch per chunk); np0 = their count
    double* defer_partial = nullptr;
    double* d_chunk_partials = nullptr;   // GPAW_CHUNK_MAX x np0
    int np0 = 0;
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    // chunks of V-cycles (solve_neutral): phi after each cycle of a chunk
    // (Jacobi: each cycle writes its result into snap[i] and never writes
    // its input; Gauss-Seidel: phi copied to snap[i] before cycle i), one
    // error slot per cycle, graphs by (length, start state)
    std::vector<double*> snap;
    double* tmp2 = nullptr;    // second level-0 sweep buffer of the Jacobi chunks
    // level-0 sweep buffers of the cycle being emitted (total = 0: phiA / phiB
    // ping-pong): sweep j of `total` writes t[(j - 1) % 2], the last writes dst
    struct {
        double* dst = nullptr;
        double* t[2] = {nullptr, nullptr};
        int done = 0, total = 0;
    } rot;
    double* d_errs = nullptr;
    double* d_sum = nullptr;   // 1024 partial sums of the charge integral
    // rho[0] holds a density gpaw_poisson_neutralize wrote (a solve may
    // then take rho = nullptr)
    bool rho_loaded = false;
    double* h_errs = nullptr;
    std::vector<hipGraphExec_t> chunk;   // index 2 * (k - 1) + state
    int use_graph = 1;
    int tuned = 0;             // launch shapes tuned (when the tuner is on)
    // V-cycles the last converged solve took and its eps: the first chunk of
    // the next solve to the same eps runs that many (0: none yet, start with 2)
    int hint = 0;
    double hint_eps = 0.0;
    int swaps = 0;             // 1 if level 0 does an odd number of sweeps per cycle
    int lc = -1;               // first level run by the fused coarse kernel (-1: none)
    gpaw::CoarsePlan* d_coarse = nullptr;
    short* d_coarse_nb = nullptr;     // neighbour tables of the fused coarse levels
    int coarse_lds = 0;        // dynamic LDS bytes of the fused kernel
    bool fuse_jacobi = true;   // fused Jacobi sweeps on levels > 0 (GPAW_AMD_FUSED_JACOBI=0: off)
    // coarse-grid correction applied by the first fused post-smoothing launch
    // (GPAW_AMD_FOLD_CORRECTION=0: a separate interpolate-and-subtract launch)
    bool fold_correction = true;
    // the residual and its restriction in one launch (GPAW_AMD_FUSED_RESTRICT=0:
    // a residual pass, then a restriction pass)
    bool fuse_restrict = true;
    // in a chunk, a V-cycle's error is summed by the next cycle's first
    // level-0 sweep (GPAW_AMD_FUSED_NORM=0: a norm pass per cycle)
    bool fuse_norm = true;
    double* norm_in = nullptr;   // partials row for the next emitted cycle's first sweep
    bool skip_norm = false;      // the next emitted cycle takes no norm pass
    gpaw::CoarsePlan h_coarse; // host copy of the coarse plan
    MixedState* mixed = nullptr;
};

void gpaw_mixed_free(MixedState* m);
