// swe_internal.hpp -- kernel argument blocks shared by the step kernels and
// the host side of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace swe {

constexpr int CHUNK = 256;  // cells per workgroup (four wave64s)

// Reduction slots of one step.  Invariant between launches: dtmin holds the
// bits of huge(1.0), counter and nref are 0.  The last workgroup of a step
// restores the invariant after publishing the results.
struct Slots {
    unsigned long long dtmin;    // running min of the CFL candidates (double bits)
    unsigned long long nref;     // running count of refinement flags
    unsigned int counter;        // workgroups done
    unsigned int pad;
    double dt;                   // dt of the next step (device dt chain)
    double time;                 // accumulated simulated time (device dt chain)
    double res_dtcfl;            // published CFL minimum of the last step
    unsigned long long res_nref; // published refinement count of the last step
};

struct StepArgs {
    const double *eta, *hu, *hv, *b;       // state at t_n, SoA, SFC order
    double *eta_out, *hu_out, *hv_out;     // state at t_{n+1}
    const int32_t *nbr;                    // [3n] (o << 2) | back-edge, -1 boundary
    const int8_t *plotter;                 // [n] forward plotter type
    const int8_t *depth;                   // [n]
    int8_t *refine;                        // [n] or null
    Slots *slots;
    double *out_dtcfl;                     // extra copy of the CFL minimum, or null
    int64_t *out_nref;                     // extra copy of the refinement count, or null
    int64_t n;
    double dt;                             // used when device_dt == 0
    double scaling, g, drytol, refine_thr;
    double courant, dt_clamp;              // device dt chain: dt' = min(clamp, courant * min)
    int dmin, dmax, bc, forward;
    int device_dt;                         // 1: read dt from slots->dt and write the next one
};

hipError_t upload_tables(const double *normal, const double *leg);
hipError_t launch_step(const StepArgs &a, int solver, hipStream_t s);

}  // namespace swe
