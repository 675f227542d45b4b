// swe_internal.hpp -- kernel argument blocks shared by the step kernels and
// the host side of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace swe {

// Chunk sizes (cells per workgroup, one thread per cell) the library is
// compiled for; the batch autotuner picks one of them.
constexpr int CHUNK_SIZES[3] = {128, 256, 512};
constexpr int DEFAULT_CHUNK = 256;

// Per-scenario results and device dt chain.  Written only by the finalize
// kernel that follows each step kernel on the stream; the step kernel reads
// dt.
struct Slots {
    double dt;                   // dt of the next step (device dt chain)
    double time;                 // simulated time (device dt chain)
    double res_dtcfl;            // CFL minimum of the last step
    long long res_nref;          // refinement count of the last step
};

// One scenario (mesh + constants) of a batch.  A single mesh is a batch of
// one.  Scenarios of equal meshes share one copy of the topology.
struct Scen {
    int64_t cell_off;    // first cell in the state arrays
    int64_t topo_off;    // first cell in the topology arrays
    int64_t n;           // cells
    int64_t chunk_off;   // first global chunk
    int64_t nchunks;     // chunks of this scenario
    double scaling, drytol, refine_thr, courant, dt_clamp;
    int32_t depth;       // uniform depth, or -1 to read depth[] per cell
    int32_t dmin, dmax;
    int32_t pad;
};

struct StepArgs {
    const double *eta, *hu, *hv, *b;       // state at t_n, SoA, SFC order
    double *eta_out, *hu_out, *hv_out;     // state at t_{n+1}
    const int32_t *nbr;                    // topology [3n] (o << 2) | back edge, -1 boundary
    const int8_t *plotter;                 // topology [n] forward plotter type
    const int8_t *depth;                   // topology [n] (used when Scen::depth < 0)
    int8_t *refine;                        // state-indexed [n] or null
    int8_t *masks;                         // state-indexed [n] limiter masks (bit 0 wetting,
                                           // bit 1 drying fix taken) or null
    const Scen *scen;                      // [n_scen], or null: one scenario, scen0
    const int32_t *chunk_scen;             // [nchunks] scenario of each chunk, or null (one scenario)
    Slots *slots;                          // [n_scen]
    const double *dt_host;                 // [n_scen] dt per scenario when device_dt == 0, or null
    double *out_dtcfl;                     // [n_scen] extra copy of the CFL minima, or null
    int64_t *out_nref;                     // [n_scen] extra copy of the refinement counts, or null
    double *part_min;                      // [nchunks] CFL minimum of each chunk
    int32_t *part_nref;                    // [nchunks] refinement count of each chunk
    int32_t n_scen;
    int64_t nchunks;
    double dt;                             // dt when device_dt == 0 and dt_host is null
    double g, inv16g;                      // inv16g = 1 / (16 g), riemannsolvers.f:556
    int bc, forward;
    int device_dt;                         // 1: read dt from slots[s].dt and write the next one
    int skip_finalize;                     // 1: nobody reads this step's CFL results
    // refinement flags from surface-elevation jumps between wet neighbours
    // of the state the step reads (refine above jump_refine, coarsen below
    // jump_coarsen) instead of volume_op's |dQ_h| criterion
    int flag_jumps;
    double jump_refine, jump_coarsen;
    // nbr holds plain neighbour cells (-1 on the boundary; swe_mesh_desc's
    // layout) instead of the packed (o << 2) | back edge: the back edge of an
    // in-chunk neighbour is looked up in the chunk's staged topology, and a
    // neighbour outside [0, n) reads as a domain boundary
    int raw_nbr;
    Scen scen0;                            // the scenario when scen is null
    // 1: toleranced arithmetic (swe_riemann_tol.hpp) for the GeoClaw solvers
    int tol;
};

// Discrete adjoint of one traversal (swe_adjoint_step.hip).
struct AdjArgs {
    const double *eta, *hu, *hv, *b;        // state at t_n
    const int8_t *masks;                    // limiter masks of the forward traversal of t_n
    const double *leta, *lhu, *lhv;         // adjoint of the state at t_{n+1}
    double *leta_out, *lhu_out, *lhv_out;   // adjoint of the state at t_n
    double *gb;                             // adjoint of b, accumulated (+=)
    const int32_t *nbr;
    const int8_t *plotter;
    const int8_t *depth;
    const Scen *scen;
    const int32_t *chunk_scen;
    int64_t nchunks;
    const double *dt_dev;                   // [n_scen] dt per scenario, or null: use dt
    double dt;
    double g, inv16g;
    int bc, forward;
};

// Remeshing of one mesh (swe_adapt.hip).  Topology in the step kernels'
// format (nbr packed (o << 2) | back edge); scratch sized for n cells.
struct AdaptArgs {
    int64_t n;
    const int8_t *depth, *plotter;
    const int32_t *nbr;
    const int8_t *flags;                   // requests -1 / 0 / 1 per cell
    const double *eta, *hu, *hv, *b;
    int dmin, dmax;
    uint32_t *slots;                       // [n] conformity words
    int64_t *off;                          // [n] tree offsets
    int64_t *base;                         // [n] prefix of the new-cell counts
    int64_t *bsum;                         // [adapt_scan_tiles(n) + 1 + 8]
    int8_t *code;                          // [n]
    int32_t *kidx;                         // [n] new index of a kept cell, -1 if it changes
    int32_t *eslot;                        // [6n] new cell-edge on each half of each old edge
                                           //      (changing cells only)
    uint32_t *status;                      // [3], zero before the first call
    uint32_t *stamp;                       // [n / 256 + 1] per tile of the conformity sweeps
    int32_t *list;                         // [n] the changing cells (written by the prefix pass)
    int64_t n_changing;                    // their number (from adapt_plan)
    int64_t *off_out;                      // [n'] tree offsets of the new mesh
    int8_t *depth_out, *plotter_out;       // [n'] new mesh
    int32_t *nbr_out;                      // [3n']
    double *eta_out, *hu_out, *hv_out, *b_out;
    Scen *scen_dev;                        // non-null: emit stores scen_new there (the
    Scen scen_new;                         // new mesh's entry of the step kernels' table)
    // bathymetry of split cells: bilinear in a raster of nx x ny values at
    // world points (x0 + i dx, y0 + j dy), values[j nx + i], at world points
    // scaling * c + offset of unit-square points c; null: the parent's b
    const double *raster;
    int64_t rnx, rny;
    double rx0, ry0, rdx, rdy, scaling, offset_x, offset_y;
    // a new cell at depth d averages the raster over min(16, raster_cap - d)
    // halvings (_ADAPT_INTEGRATE); raster_cap = min(dmax + 3, data_depth + 3)
    int raster_cap;
};

// bsum holds adapt_scan_tiles(n) + 1 scan words and 8 result words.
int64_t adapt_scan_tiles(int64_t n);
struct AdaptPlan {
    int sweeps;          // conformity sweeps run, -1 if the fixpoint was not reached in max_sweeps
    int last_change;     // the last of them that changed a word (0: none did)
    bool prefix_done;    // base and kidx are written (prefix_ahead)
    int64_t n_new;       // cells of the new mesh
    int64_t tree_total;  // covered part of the two root trees (2^61 when the mesh is whole)
    uint32_t status;     // bit 0: an inconsistent merge; bit 1: the mesh changes
    int64_t n_changing;  // cells that split or merge
};
// Marks the edges and sweeps to the conformity fixpoint (first_batch sweeps,
// then 2 at a time; the first sweep starts from the requests), decides every
// cell (code) and sums the new-cell counts per scan tile; the count pass
// confirms the fixpoint.  offsets_valid: a.off still holds this mesh's tree
// offsets.  prefix_ahead: launch the prefix of the counts behind every count
// pass, before the host knows whether the mesh changes (it runs while the
// host reads).  epoch: a number that differs from the last few calls' (it
// tags their sweep stamps, so that no pass has to clear them).  host: 8
// words of coherent pinned memory for the results, host_dev its device
// address; *seq numbers the results (incremented per batch).  Waits for the
// device once per batch.  adapt_emit then writes the prefix of the counts
// (base; its high word counts the cells that change) and the new mesh.
hipError_t adapt_plan(const AdaptArgs &a, hipStream_t s, int first_batch, bool offsets_valid,
                      bool prefix_ahead, int max_sweeps, uint32_t epoch, int64_t *host,
                      int64_t *host_dev, int64_t *seq, AdaptPlan *plan);
// Refinement flags from surface-elevation jumps to wet neighbours.
hipError_t flag_eta_jumps(int64_t n, const double *eta, const double *b, const int32_t *nbr,
                          const int8_t *depth, double drytol, double refine_tol,
                          double coarsen_tol, int dmin, int dmax, int8_t *flags, hipStream_t s);
// Transpose of a remeshing's state transfer (code, base: its record).
hipError_t transfer_adjoint(int64_t n_old, const int8_t *code, const int64_t *base, int b_sampled,
                            const double *le, const double *lu, const double *lv, const double *gb,
                            double *le_o, double *lu_o, double *lv_o, double *gb_o,
                            hipStream_t s);
// The record of a remeshing that changed nothing.
hipError_t identity_record(int64_t n, int8_t *code, int64_t *base, hipStream_t s);
// out[c] = kidx[c] where the cell is kept, else base[c] (base holds the
// changing cells' prefixes only)
hipError_t record_base(int64_t n, const int32_t *kidx, const int64_t *base, int64_t *out,
                       hipStream_t s);
// Writes the new mesh and state (outputs sized for n_new).
// prefix_done: adapt_plan has launched the prefix pass already.
hipError_t adapt_emit(const AdaptArgs &a, hipStream_t s, bool prefix_done);

hipError_t upload_tables(const double *normal, const double *leg);
hipError_t upload_tables_adjoint(const double *normal, const double *leg);
hipError_t launch_adjoint(const AdjArgs &a, int solver, int chunk, int64_t grid_blocks,
                          hipStream_t s);
// Launches the step kernel and, after it on the same stream, the finalize
// kernel (per-scenario reduction of the chunk partials, results, dt chain).
// grid_blocks: number of workgroups (multiple of 8) or 0 for one per chunk.
hipError_t launch_step(const StepArgs &a, int solver, int chunk, int64_t grid_blocks,
                       hipStream_t s);
// The step of a single mesh whose nbr holds plain neighbour cells
// (StepArgs::raw_nbr), chunk DEFAULT_CHUNK, one workgroup per chunk, flags
// by volume_op's criterion; then the finalize kernel.
hipError_t launch_step_raw(const StepArgs &a, int solver, hipStream_t s);

}  // namespace swe
