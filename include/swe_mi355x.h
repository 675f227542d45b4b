/*
 * swe_mi355x.h -- C ABI of the MI355X shallow-water (SWE) Euler step.
 *
 * A drop-in for the per-step edge/cell traversal of sam(oa)^2's SWE scenario:
 * one call replaces `swe%euler%traverse(grid)` (SWE.f90:393), i.e. the
 * operators of SWE_euler_timestep.f90:36-54 run over every cell and edge of
 * the mesh (flux = GeoClaw augmented Riemann solver on every edge, cell
 * update with the wet/dry limiter, CFL reduction).  Between steps, swe_adapt
 * replaces the adaption traversal (SWE.f90:387-390) on the device.  Plain pointers and sizes
 * only; every entry point returns SWE_OK (0) or a negative status and never
 * aborts.  The text of the last error of the calling thread is available from
 * swe_last_error().
 *
 * Data conventions (all arrays in forward Sierpinski SFC order, as
 * SFC_node_traversal.f90:90-145 generates them):
 *   eta[n]     surface elevation (the reference's Q%h), metres
 *   hu[n],hv[n] momentum in world coordinates (Q%p)
 *   b[n]       bathymetry (Q%b)
 *   depth[n]   refinement depth d of each cell (fine_triangle%i_depth)
 *   plotter[n] plotter-grammar type in forward orientation, -8..8 without 0
 *              (fine_triangle%i_plotter_type)
 *   nbr[3n]    neighbour across edge k = 0,1,2 (reference edge index k+1,
 *              edge i opposite node i, SFC_data_types.f90:660-666), -1 on the
 *              domain boundary; consecutive cells of the curve share an edge
 *              (swe_create rejects a mesh where they do not)
 *
 * Ownership: the caller owns host arrays and may free them when a call
 * returns; the library owns its device buffers.  A handle is not re-entrant:
 * calls on one handle must not overlap.
 */
#ifndef SWE_MI355X_H
#define SWE_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWE_OK 0
#define SWE_ERR_ARG (-1)      /* invalid argument (null pointer, bad size, bad mesh) */
#define SWE_ERR_HIP (-2)      /* a HIP runtime call failed */
#define SWE_ERR_NOMEM (-3)    /* device or host allocation failed */
#define SWE_ERR_STATE (-4)    /* call not valid in the handle's state */

#define SWE_SOLVER_AUG_RIEMANN 0 /* riemannsolvers.f:2-258 (flux_solver=aug_riemann) */
#define SWE_SOLVER_FWAVE 1       /* riemannsolvers.f:446-500 (flux_solver=fwave) */
/* compute_lf_flux, SWE_euler_timestep.f90:359-446: global and local
 * Lax-Friedrichs (flux_solver=lf, llf; a flat bed b = 0) and their
 * well-balanced variants with bathymetry and wet/dry walls (lfbath,
 * llfbath).  Forward steps only: the adjoint takes the GeoClaw solvers. */
#define SWE_SOLVER_LF 2
#define SWE_SOLVER_LLF 3
#define SWE_SOLVER_LF_BATH 4
#define SWE_SOLVER_LLF_BATH 5

#define SWE_BC_OUTFLOW 0 /* ghost = copy of the cell (SWE_euler_timestep.f90:248) */
#define SWE_BC_WALL 1    /* ghost momentum mirrored, p - 2 (p.n) n */

typedef struct swe_handle swe_handle;

/* Mesh and scenario constants; the fields a `t_swe` run takes from `cfg`
 * (Config.f90:243,284, SWE.f90:169-170) and from the grid. */
typedef struct swe_mesh_desc {
    int64_t n_cells;
    const int8_t *depth;    /* [n_cells] */
    const int8_t *plotter;  /* [n_cells] */
    const int32_t *nbr;     /* [3*n_cells] */
    double scaling;         /* cfg%scaling: edge length of the unit square in metres */
    double g;               /* gravitational constant (SWE_data_types.f90:34: 9.80665), in [1e-10, 1e10] */
    double dry_tolerance;   /* cfg%dry_tolerance, >= 1e-100 */
    double courant;         /* cfg%courant_number (applied by swe_run only) */
    int32_t min_depth;      /* cfg%i_min_depth, for the refinement flags */
    int32_t max_depth;      /* cfg%i_max_depth */
    int32_t solver;         /* SWE_SOLVER_* */
    int32_t boundary;       /* SWE_BC_* */
} swe_mesh_desc;

/* Creates a handle: validates and uploads the mesh topology.  Replaces the
 * grid set-up the traversal template performs per scenario
 * (SFC_generic_traversal_ringbuffer.f90:252-285). */
int swe_create(const swe_mesh_desc *desc, swe_handle **out);

/* Copies a host state in.  b is constant during steps. */
int swe_upload(swe_handle *h, const double *eta, const double *hu, const double *hv,
               const double *b);

/* Copies the current state out (any pointer may be NULL to skip it). */
int swe_download(swe_handle *h, double *eta, double *hu, double *hv, double *b);

/* One Euler traversal: replaces swe%euler%traverse(grid) (SWE.f90:393).
 *   dt       grid%r_dt after pre_traversal_grid_op (SWE_euler_timestep.f90:89-109)
 *   forward  1 if the traversal visits cells in forward SFC order, 0 if backward
 *            (the reference flips the direction after every traversal,
 *            SFC_generic_traversal_ringbuffer.f90:420); decides which cell is
 *            the left state of each edge and the order of the update sum.
 *   dt_cfl   out: min over cells of vol / sum(len_i * s_i) (the reduced
 *            grid%r_dt_new before the courant factor, :118)
 *   n_refine out (nullable): number of cells flagged for refinement (:117)
 *   refine   out (nullable): refinement flag per cell, -1/0/1 (:336-344)
 * The host keeps t += dt and dt = courant * dt_cfl, as post_traversal_grid_op
 * (:111-131) does.  Synchronous. */
int swe_step(swe_handle *h, double dt, int forward, double *dt_cfl, int64_t *n_refine,
             int8_t *refine);

/* Runs n_steps Euler traversals on the device without host round trips: the
 * dt chain of post_traversal_grid_op / pre_traversal_grid_op is evaluated on
 * the device (dt_{k+1} = min(dt_clamp, courant * dt_cfl_k); dt_clamp <= 0
 * disables the clamp).  The direction starts at `forward` and flips after
 * every step.  *time is the simulated time before the steps on entry and is
 * advanced by each step's dt on the device (t = t + dt, in the reference's
 * order); *dt_next receives the dt of the next step.  The steps are captured
 * once into a HIP graph and replayed.
 * dt0 <= 0 continues the device chain (dt and time) where the last swe_run
 * left it; *time is then not read.  With time and dt_next both NULL the call
 * returns without waiting for the steps (time is required when dt0 > 0). */
int swe_run(swe_handle *h, int64_t n_steps, double dt0, int forward, double dt_clamp,
            double *time, double *dt_next);

/* Device-pointer step for callers that own device memory (the PyTorch
 * binding): reads eta/hu/hv/b at the given device addresses, writes
 * eta_out/hu_out/hv_out (must not alias the inputs), on `stream`
 * (a hipStream_t, NULL = the library's stream).  dt_cfl_dev and
 * n_refine_dev (nullable) are device scalars that receive the step's CFL
 * minimum and refinement count; refine_dev (nullable) receives the flags.
 * Asynchronous; calls on one handle must not overlap, whatever their streams. */
int swe_step_device(swe_handle *h, const double *eta, const double *hu, const double *hv,
                    const double *b, double *eta_out, double *hu_out, double *hv_out,
                    double dt, int forward, double *dt_cfl_dev, int64_t *n_refine_dev,
                    int8_t *refine_dev, void *stream);

/* ---- The step without device allocations -----------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): swe_step_device without a handle.  The caller passes the
 * mesh, the state and the outputs as device addresses; the step's scratch
 * (the CFL minima and refinement counts of its 256-cell chunks and the
 * finalize kernel's result record) lives in a fixed device array of the
 * library, 768.25 KiB per device, held until the process ends.  Limit:
 * n_cells <= 16,777,216 (65,536 chunks); larger meshes return SWE_ERR_ARG.
 * The call allocates nothing (no hipMalloc, no host allocation) and is
 * asynchronous on `stream` (a hipStream_t; NULL = the null stream).  The
 * first call on a device also looks up the scratch's address and uploads the
 * geometry tables there (synchronous copies): make it outside graph capture.
 * Every call on a device shares that device's scratch: calls that overlap on
 * one device (other streams, other threads) race; calls in order on one
 * stream do not.  Nothing in the scratch must start at zero.  The mesh is not
 * validated (check it once with swe_create): nbr holds swe_mesh_desc's plain
 * neighbour cells, and an entry outside [0, n_cells) reads as a domain
 * boundary; depth and plotter must be in their documented ranges.  The
 * chunk size is the library's default (swe_chunk_cells), one workgroup per
 * chunk; results equal swe_step_device's, which does not depend on its
 * launch shape. */
typedef struct swe_device_mesh {
    int64_t n_cells;
    const int8_t *depth;    /* device [n_cells] */
    const int8_t *plotter;  /* device [n_cells] */
    const int32_t *nbr;     /* device [3*n_cells], neighbour cell or -1 */
    double scaling;         /* as swe_mesh_desc */
    double g;
    double dry_tolerance;
    int32_t min_depth, max_depth; /* for the refinement flags */
    int32_t solver;         /* SWE_SOLVER_* */
    int32_t boundary;       /* SWE_BC_* */
} swe_device_mesh;

int swe_step_device_noalloc(const swe_device_mesh *mesh, const double *eta, const double *hu,
                            const double *hv, const double *b, double *eta_out, double *hu_out,
                            double *hv_out, double dt, int forward, double *dt_cfl_dev,
                            int64_t *n_refine_dev, int8_t *refine_dev, void *stream);

/* ---- Discrete adjoint ------------------------------------------------------
 * No counterpart in the reference (the reference has no adjoint).  For the
 * traversal Q^{n+1} = F(Q^n, b; dt_n) with dt_n and the wet/dry limiter
 * decisions frozen from the forward pass, swe_adjoint_step_device computes
 * lam^n = (dF/dQ^n)^T lam^{n+1} and adds (dF/db)^T lam^{n+1} to gb.
 * All pointers are device addresses; asynchronous on `stream`.
 *
 * swe_forward_device: one traversal on caller buffers with dt read from
 * device memory (*dt_dev), the CFL minimum written to *dt_cfl_dev (nullable)
 * and the limiter masks (bit 0 wetting fix, bit 1 drying fix of
 * cell_update_op, SWE_euler_timestep.f90:271-283) to masks_dev (nullable),
 * for the recomputation sweeps of a checkpointed adjoint. */
int swe_forward_device(swe_handle *h, const double *eta, const double *hu, const double *hv,
                       const double *b, double *eta_out, double *hu_out, double *hv_out,
                       const double *dt_dev, int forward, double *dt_cfl_dev, int8_t *masks_dev,
                       void *stream);
int swe_adjoint_step_device(swe_handle *h, const double *eta, const double *hu, const double *hv,
                            const double *b, const int8_t *masks, const double *leta,
                            const double *lhu, const double *lhv, double *leta_out,
                            double *lhu_out, double *lhv_out, double *gb, const double *dt_dev,
                            int forward, void *stream);

/* ---- remeshing --------------------------------------------------------- */

/* Replaces one call of swe%adaption%traverse(grid) (SWE.f90:387-390): the
 * conformity check of Conformity/Conformity.f90:172-208 (mark_edges and
 * mark_edges_integrity, :1244-1317) on the refinement requests, then the
 * adaptive traversal of SFC_generic_adaptive_traversal.f90 with the state
 * transfer of SWE_adapt.f90:95-169.  A cell whose hypotenuse is marked
 * splits (a marked leg splits that child again); a pair of sibling cells
 * whose legs are marked merges; the others stay.  A split copies eta, hu, hv
 * and b to the children; a merge averages all four, 0.5 (Q1 + Q2).  (The
 * reference evaluates the scenario's bathymetry for a split cell; for a flat
 * or cell-wise constant bed that is the parent's value.)
 *
 * refine: the requests (-1 coarsen, 0 keep, 1 refine) per cell in curve
 * order, as swe_step writes them; NULL takes the flags the handle's last
 * step left on the device (swe_keep_refinement_flags must be on).  Requests
 * beyond [min_depth, max_depth] are dropped, as volume_op never issues them.
 * The mesh must be whole (the leaf sequence of the two root triangles).
 * Writes the new cell count to *n_cells and the number of conformity sweeps
 * to *sweeps (both nullable).  The traversal direction of the reference flips
 * once per adaption: the caller flips `forward` once more.  Single-mesh
 * handles only.  Device memory grows as the mesh does; it is not returned
 * when the mesh shrinks. */
int swe_adapt(swe_handle *h, const int8_t *refine, int64_t *n_cells, int *sweeps);

/* Bathymetry of the cells that swe_adapt splits, in place of the parent's
 * value, as get_bathymetry_at_element (SWE_initialize.f90:120-147) computes
 * it with gridded data (the ASAGI path, SWE_initialize.f90:149-187) in the
 * reference's _ADAPT_INTEGRATE build.  The raster is bilinear in nx x ny
 * values over the world points (x0 + i dx, y0 + j dy), values[j * nx + i],
 * clamped to its last cell; a unit-square point x lies at world
 * scaling * x + (offset_x, offset_y) (cfg%scaling, cfg%offset).  A child at
 * depth d averages it over L = min(16, min(max_depth + 3, data_depth + 3) - d)
 * halvings, data_depth = nint(log2(scaling^2 / (dx dy))): each level
 * 0.5 (value of the half (x1, m, x2) + value of the half (x2, m, x3)),
 * m = 0.5 (x1 + x3), and at the bottom the raster at the centroid
 * (x1 + x2 + x3) / 3 (refine_2D_recursive, :106-116); L <= 0 is that
 * centroid sample alone.  The library copies the values; NULL removes the
 * raster. */
int swe_set_bathymetry_raster(swe_handle *h, int64_t nx, int64_t ny, double x0, double y0,
                              double dx, double dy, double offset_x, double offset_y,
                              const double *values);

/* Refinement flags from the gradient of the surface elevation, for runs that
 * refine on |grad eta| rather than on volume_op's |dQ_h| criterion
 * (SWE_euler_timestep.f90:336-344); no counterpart in the reference.  For a
 * wet cell (eta - b > dry_tolerance), j = the largest |eta - eta_n| over its
 * wet neighbours n; a dry cell has j = 0.  The flag is 1 if j > refine_tol
 * and depth < max_depth, -1 if j < coarsen_tol and depth > min_depth, else
 * 0.  Replaces the device flags that swe_adapt(h, NULL, ...) reads.
 * Asynchronous on the handle's stream. */
int swe_flag_eta_jumps(swe_handle *h, double refine_tol, double coarsen_tol);

/* on != 0: every step of swe_step and swe_run leaves its refinement flags on
 * the device for swe_adapt(h, NULL, ...), with volume_op's |dQ_h| criterion
 * (SWE_euler_timestep.f90:336-344).  Either value turns the elevation-jump
 * criterion of swe_keep_eta_jump_flags off. */
int swe_keep_refinement_flags(swe_handle *h, int on);

/* Every step of swe_step and swe_run leaves refinement flags on the device
 * for swe_adapt(h, NULL, ...) that follow the criterion of
 * swe_flag_eta_jumps (refine_tol, coarsen_tol >= 0), evaluated on the state
 * the step reads, as volume_op evaluates its criterion on the state of its
 * traversal; no counterpart in the reference.  The traversal finds both
 * states of every edge anyway, so no separate pass over the mesh is made.
 * The flags of step k drive the remeshing before step k + 1, as in the
 * reference's time loop (SWE.f90:385-393).  Single-mesh handles only;
 * swe_step's refine output returns these flags too. */
int swe_keep_eta_jump_flags(swe_handle *h, double refine_tol, double coarsen_tol);

/* The record of the last swe_adapt, for the adjoint of a run with
 * remeshing (no counterpart in the reference): code[n_old] (0 keep, 1 first
 * and 2 second of a merging pair, 4 + pattern for a split into 2-4 cells)
 * and base[n_old] (the first new cell of each old cell in its low 32 bits),
 * device addresses of the caller's; *n_old the old cell count, *b_sampled
 * whether split cells sampled the bathymetry raster.  NULL code and base
 * return the sizes only.  Asynchronous on `stream` (NULL: the handle's). */
int swe_adapt_record_device(swe_handle *h, int8_t *code, int64_t *base, int64_t *n_old,
                            int *b_sampled, void *stream);

/* Transpose of a remeshing's state transfer (no counterpart in the
 * reference): from the adjoint of the new state (leta, lhu, lhv, gb over
 * the new cells) to that of the old state (over the n_old old cells): a
 * split's children sum into the parent, each cell of a merged pair takes
 * half of the parent's.  gb_out accumulates (+=); split cells that sampled
 * the raster (b_sampled) pass nothing to the parent's b.  Device pointers,
 * asynchronous on `stream` (a hipStream_t). */
int swe_transfer_adjoint_device(int64_t n_old, const int8_t *code, const int64_t *base,
                                int b_sampled, const double *leta, const double *lhu,
                                const double *lhv, const double *gb, double *leta_out,
                                double *lhu_out, double *lhv_out, double *gb_out, void *stream);

/* Cells of the handle's current mesh (changes with swe_adapt); -1 for NULL. */
int64_t swe_cells(swe_handle *h);

/* The current mesh in swe_mesh_desc's layout (each argument nullable):
 * depth[n], plotter[n], nbr[3n] (neighbour cell or -1). */
int swe_download_mesh(swe_handle *h, int8_t *depth, int8_t *plotter, int32_t *nbr);

/* Launch subsequent work of this handle on `stream` (a hipStream_t; NULL
 * restores the library's own stream), e.g. to time it with events the caller
 * records on that stream.  Waits for work already queued. */
int swe_set_stream(swe_handle *h, void *stream);

/* Launch shape of the handle's step kernel, as swe_batch_set_tuning below:
 * chunk_cells in {128, 256, 512}, grid_blocks 0 or a multiple of 8.  No
 * counterpart in the reference; results do not depend on it. */
int swe_set_tuning(swe_handle *h, int chunk_cells, int64_t grid_blocks);
int swe_get_tuning(swe_handle *h, int *chunk_cells, int64_t *grid_blocks);

/* Arithmetic of the step kernels (no counterpart in the reference).
 * SWE_ARITH_EXACT (the default): every operation rounds as the reference's
 * build with -ffp-contract=off rounds it; results are bit-identical to it.
 * SWE_ARITH_TOLERANCED: the GeoClaw solvers (SWE_SOLVER_AUG_RIEMANN,
 * SWE_SOLVER_FWAVE) share reciprocals of repeated denominators, take
 * hardware reciprocal and square-root estimates refined by Newton steps and
 * solve each edge from per-cell depths and velocities, with the exact path's
 * left and right sides and no fused multiply-adds; one step stays within a
 * few ulps of the terms it sums (tested against the exact path at 1e-12 of
 * the largest field value per step and 1e-9 after 1000 steps), and
 * refinement flags can differ where a criterion sits within rounding of its
 * threshold.  The Lax-Friedrichs
 * solvers and the device-pointer sweeps of the adjoint (swe_forward_device,
 * swe_batch_forward_device) always round exactly.  Applies to swe_step,
 * swe_run and swe_step_device (and their batch forms) from the next call. */
#define SWE_ARITH_EXACT 0
#define SWE_ARITH_TOLERANCED 1
int swe_set_arithmetic(swe_handle *h, int mode);

/* Releases the handle and its device memory. NULL is accepted. */
int swe_destroy(swe_handle *h);

/* Copies the last error message of the calling thread into buf (always NUL
 * terminated when size > 0); returns its full length. */
size_t swe_last_error(char *buf, size_t size);

/* Default number of cells a workgroup processes (the LDS chunk). */
int swe_chunk_cells(void);

/* ---- Batched ensembles ----------------------------------------------------
 * Many independent scenarios (meshes of uneven size, each with its own
 * scaling, dry tolerance and dt chain) advanced together by one kernel launch
 * per step through a per-scenario cell-offset table.  No counterpart in the
 * reference (one grid per process, SFC_traversal.f90:38-139): each scenario
 * behaves exactly as a single-mesh handle would.  Scenarios whose desc points
 * at the same depth/plotter/nbr arrays share one device copy of the topology.
 * solver, boundary and g must be equal across scenarios.  All scenarios step
 * in the same traversal direction. */
typedef struct swe_batch swe_batch;

int swe_batch_create(int32_t n_scenarios, const swe_mesh_desc *descs, swe_batch **out);
int swe_batch_upload(swe_batch *b, int32_t scenario, const double *eta, const double *hu,
                     const double *hv, const double *bathymetry);
int swe_batch_download(swe_batch *b, int32_t scenario, double *eta, double *hu, double *hv,
                       double *bathymetry);
/* One traversal of every scenario: dt[s] in, dt_cfl[s] out (and n_refine[s],
 * nullable), as swe_step per scenario.  Synchronous. */
int swe_batch_step(swe_batch *b, const double *dt, int forward, double *dt_cfl,
                   int64_t *n_refine);
/* n_steps traversals of every scenario with per-scenario device dt chains, as
 * swe_run per scenario: dt0[s] and time[s] (in/out), dt_next[s] out.  The
 * steps are replayed from a captured HIP graph. */
int swe_batch_run(swe_batch *b, int64_t n_steps, const double *dt0, int forward,
                  double dt_clamp, double *time, double *dt_next);
/* Launch shape: chunk_cells in {128, 256, 512} (cells per workgroup, one LDS
 * tile), grid_blocks 0 (one workgroup per chunk) or a multiple of 8 (a
 * persistent grid that strides over the chunks).  Results do not depend on
 * it; the autotuner (samoa_amd/autotune.py) picks the fastest. */
int swe_batch_set_tuning(swe_batch *b, int chunk_cells, int64_t grid_blocks);
int swe_batch_get_tuning(swe_batch *b, int *chunk_cells, int64_t *grid_blocks);
int swe_batch_set_stream(swe_batch *b, void *stream);
/* As swe_set_arithmetic, for every scenario of the batch. */
int swe_batch_set_arithmetic(swe_batch *b, int mode);
/* Total cells of all scenarios, or -1 for a NULL batch. */
int64_t swe_batch_cells(swe_batch *b);
/* First cell of a scenario in the concatenated state arrays, or -1. */
int64_t swe_batch_offset(swe_batch *b, int32_t scenario);
/* Batched forms of swe_forward_device / swe_adjoint_step_device over the
 * concatenated state arrays of all scenarios; dt_dev[s] per scenario;
 * dt_cfl_dev[s] (nullable) per scenario. */
int swe_batch_forward_device(swe_batch *bt, const double *eta, const double *hu, const double *hv,
                             const double *b, double *eta_out, double *hu_out, double *hv_out,
                             const double *dt_dev, int forward, double *dt_cfl_dev,
                             int8_t *masks_dev, void *stream);
int swe_batch_adjoint_device(swe_batch *bt, const double *eta, const double *hu, const double *hv,
                             const double *b, const int8_t *masks, const double *leta,
                             const double *lhu, const double *lhv, double *leta_out,
                             double *lhu_out, double *lhv_out, double *gb, const double *dt_dev,
                             int forward, void *stream);
int swe_batch_destroy(swe_batch *b);

#ifdef __cplusplus
}
#endif
#endif
