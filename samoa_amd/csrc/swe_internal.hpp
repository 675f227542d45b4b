// swe_internal.hpp -- kernel argument blocks shared by the step kernels and
// the host side of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace swe {

constexpr int CHUNK = 256;  // cells per workgroup (four wave64s)

struct StepArgs {
    const double *eta, *hu, *hv, *b;       // state at t_n, SoA, SFC order
    double *eta_out, *hu_out, *hv_out;     // state at t_{n+1}
    const int32_t *nbr;                    // [3n] (o << 2) | back-edge, -1 boundary
    const int8_t *plotter;                 // [n] forward plotter type
    const int8_t *depth;                   // [n]
    int8_t *refine;                        // [n] or null
    unsigned long long *dtmin;             // device min slot (bit pattern of a double)
    unsigned long long *nref;              // device refinement counter or null
    const double *dt_dev;                  // if set, dt is read from here
    int64_t n;
    double dt;
    double scaling, g, drytol, refine_thr;
    int dmin, dmax, bc, forward;
};

struct FinalizeArgs {
    double *dt;                  // in: dt used by the step; out: dt of the next step
    double *time;                // accumulated simulated time
    unsigned long long *dtmin;   // reduced CFL minimum; reset afterwards
    double courant, dt_clamp;
};

hipError_t upload_tables(const double *normal, const double *leg);
hipError_t launch_step(const StepArgs &a, int solver, hipStream_t s);
hipError_t launch_finalize(const FinalizeArgs &f, hipStream_t s);
hipError_t launch_reset(unsigned long long *dtmin, unsigned long long *nref, hipStream_t s);

}  // namespace swe
