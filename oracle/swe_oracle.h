/*
 * swe_oracle.h -- CPU restatement of the sam(oa)^2 SWE Euler step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in samoa_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the sam(oa)^2 source tree).  Arithmetic follows the reference's operation
 * order so that, compiled with -ffp-contract=off, results are bit-identical to
 * the reference built with the same flag.
 *
 * Conventions (match the HIP library):
 *   state per cell: eta (the reference's Q%h, the surface elevation), hu, hv, b
 *   cells in forward Sierpinski order; edge k (0,1,2) is the reference edge
 *   index k+1 in forward orientation (edge i opposite node i, edge 2 the
 *   hypotenuse); nbr[3*c+k] = neighbour cell or -1 on the domain boundary.
 */
#ifndef SWE_ORACLE_H
#define SWE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWO_G 9.80665
#define SWO_BC_OUTFLOW 0
#define SWO_BC_WALL 1
#define SWO_SOLVER_AUG 0
#define SWO_SOLVER_FWAVE 1
/* compute_lf_flux (SWE_euler_timestep.f90:359-446) under _LF_FLUX, _LLF_FLUX,
 * _LF_BATH_FLUX, _LLF_BATH_FLUX */
#define SWO_SOLVER_LF 2
#define SWO_SOLVER_LLF 3
#define SWO_SOLVER_LF_BATH 4
#define SWO_SOLVER_LLF_BATH 5

/* l(k) = sqrt(0.5)**k as flang folds it, k = -1..62, stored at index k+1. */
void swo_leg_sizes(double out[64]);

/* riemannsolvers.f:507-618 (maxiter = 1). */
void swo_riemanntype(double hL, double hR, double uL, double uR, double drytol,
                     double g, double *hm, double *s1m, double *s2m,
                     int *rare1, int *rare2);

/* c_bind_riemannsolvers.f90:16-176 with riemann_aug_JCP (riemannsolvers.f:2-258)
 * or riemann_fwave (riemannsolvers.f:446-500). */
void swo_geoclaw(int solver, double hL, double hR, double huL, double huR,
                 double hvL, double hvR, double bL, double bR, double drytol,
                 double g, double netL[3], double netR[3], double *maxs);

/* SWE_euler_timestep.f90:449-481: QL/QR = (eta, hu, hv, b), fL/fR = (h, hu, hv, s). */
void swo_flux(int solver, const double n[2], const double QL[4], const double QR[4],
              double drytol, double g, double fL[4], double fR[4]);

/* SFC_data_types.f90:619-700: outward unit normals of edges 1..3 of plotter type p. */
void swo_normals(int p, double n[3][2]);

/* Uniform Sierpinski mesh, SFC_node_traversal.f90:90-145,252-292.
 * Writes plotter[N], coords[6N] (node1,node2,node3 of each cell, unit square),
 * nbr[3N]. Returns N = 2^(depth+1), or -1 on allocation failure. */
int64_t swo_mesh_uniform(int depth, int8_t *plotter, double *coords, int32_t *nbr);

/* One Euler traversal, SWE_euler_timestep.f90:89-140,231-355 with the edge
 * scheduling of Tools_adaptive_traversal.f90:472-742.  Reads eta/hu/hv/b,
 * writes eta_out/hu_out/hv_out (may not alias the inputs) and refine[N].
 * Returns the minimum CFL candidate (dt_new before the courant factor) and
 * the number of refinement requests. */
double swo_step(int64_t n, const int8_t *depth, const int8_t *plotter,
                const int32_t *nbr, double scaling, double g, double drytol,
                int dmin, int dmax, int bc, int solver, int forward, double dt,
                const double *eta, const double *hu, const double *hv,
                const double *b, double *eta_out, double *hu_out,
                double *hv_out, int8_t *refine, int64_t *n_refine);

/* Initial dt estimate, SWE_initialize.f90:274-336 + :233-241 (unscaled; the
 * courant factor is applied). */
double swo_initial_dt(int64_t n, const int8_t *depth, const double *eta,
                      const double *hu, const double *hv, const double *b,
                      double g, double courant);

/* Multi-threaded variant of swo_step for CPU timing (OpenMP over cells).
 * Bit-identical to swo_step. */
double swo_step_omp(int64_t n, const int8_t *depth, const int8_t *plotter,
                    const int32_t *nbr, double scaling, double g, double drytol,
                    int dmin, int dmax, int bc, int solver, int forward, double dt,
                    const double *eta, const double *hu, const double *hv,
                    const double *b, double *eta_out, double *hu_out,
                    double *hv_out, int8_t *refine, int64_t *n_refine);

/* One remeshing step (swe_oracle_adapt.c): conformity of the refinement
 * requests flags[n] (-1/0/1, Conformity.f90:172-208,1244-1317), split and
 * merge in curve order (SFC_generic_adaptive_traversal.f90) and the state
 * transfer of SWE_adapt.f90:95-169.  coords[6n] (corners in the unit square)
 * may be NULL; with coords, coords_out and nbr_out the new corners and the new
 * neighbour table (matched through edge midpoints) are written too.  Writes
 * at most cap cells; returns the new cell count, or a negative code (-2 out of
 * memory, -3 asymmetric neighbours, -4 cap too small, -5 inconsistent merge).
 * sweeps (nullable): conformity traversals done, the reference's count. */
int64_t swo_adapt(int64_t n, const int8_t *depth, const int8_t *plotter, const int32_t *nbr,
                  const double *coords, const int8_t *flags, int dmin, int dmax,
                  const double *eta, const double *hu, const double *hv, const double *b,
                  int64_t cap, int8_t *depth_out, int8_t *plotter_out, int32_t *nbr_out,
                  double *coords_out, double *eta_out, double *hu_out, double *hv_out,
                  double *b_out, int *sweeps);

#ifdef __cplusplus
}
#endif
#endif
