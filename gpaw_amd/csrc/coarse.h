// Plan of the fused coarse-level V-cycle kernel (device memory, uploaded
// once per Poisson plan).
#pragma once
#include "transfer_common.h"

namespace gpaw {

#define GPAW_MAX_LEVELS 8

struct CoarseLevel {
    Geom g;
    int off;              // LDS mode: offset (doubles) of this level's 4 arrays
    int nb_off;           // offset of this level's neighbour table in the plan's table
    double c[7];          // 7-point star: centre, z+, z-, y+, y-, x+, x-
    double w;             // Jacobi weight
    int pre, post;        // smoothing counts
    double* phiA;         // phi ping-pong pair (A holds phi at cycle start)
    double* phiB;
    double* rho;
    double* res;
    TrArgs rst;           // restriction from this level (input = this level)
    TrArgs itp;           // interpolation into this level (input = next level)
};

struct CoarsePlan {
    int nlev;
    int lds;              // 1: all level arrays live in LDS (lds_doubles total)
    int lds_doubles;
    int nb_words;         // neighbour tables: 4-byte words (LDS mode: after the arrays)
    // per level and point, the flat indices of its six neighbours in the
    // 7-point operator's summation order (z+, z-, y+, y-, x+, x-), -1 outside
    // a zero boundary: built on the host once per plan
    const short* nbtab;
    CoarseLevel lev[GPAW_MAX_LEVELS];
};

// Arguments of the single-level kernel (a plan of one level): its post
// sweeps from phi = 0, the result stored in `out` (phiA after an even
// number of sweeps, phiB after an odd one, where single sweeps leave it).
struct CoarseOne {
    Geom g;
    double c[7];
    double w;
    int sweeps;
    const void* rho;   // double, or float for the FP64 solve inside an FP32 V-cycle
    void* out;
};

TrArgs make_trargs(const Transfer& tr);
// neighbour table of one level (6 entries per point, see CoarsePlan::nbtab)
void coarse_neighbour_table(const Geom& g, short* tab);
// h_plan: the host copy of *d_plan (a one-level plan takes the single-level
// kernel, whose arguments come from it)
hipError_t launch_coarse_vcycle(const CoarsePlan* d_plan, const CoarsePlan& h_plan, int lds_bytes,
                                hipStream_t st);
hipError_t coarse_set_lds_limit(int bytes);
// FP64 coarse solve inside an FP32 V-cycle: first-level rho / phi are float
// (LDS mode only)
hipError_t launch_coarse_vcycle_f32(const CoarsePlan* d_plan, const CoarsePlan& h_plan,
                                    int lds_bytes, hipStream_t st);

}  // namespace gpaw
